// Cantera-YAML subset parser and mechanism builder.
//
// The subset is what Cantera's own converters (ck2yaml, ctml2yaml) write:
// block mappings and sequences by indentation, flow mappings/sequences that
// may span lines, `|`/`|-` block scalars (skipped), `#` comments and CRLF line
// endings (both mechanism files in the reference are CRLF).
#include "mechanism.hpp"

#include <cctype>
#include <cmath>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>

namespace hf {
namespace {

struct Node {
    enum Type { SCALAR, MAP, SEQ } type = SCALAR;
    std::string text;
    std::vector<std::pair<std::string, Node>> map;
    std::vector<Node> seq;

    const Node* get(const std::string& key) const {
        for (auto& kv : map)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::string unquote(const std::string& s) {
    std::string t = trim(s);
    if (t.size() >= 2 && ((t.front() == '\'' && t.back() == '\'') || (t.front() == '"' && t.back() == '"')))
        return t.substr(1, t.size() - 2);
    return t;
}

// Strip a trailing comment: '#' preceded by whitespace (or at line start),
// outside quotes.
std::string strip_comment(const std::string& s) {
    bool sq = false, dq = false;
    for (size_t i = 0; i < s.size(); ++i) {
        char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (c == '#' && !sq && !dq && (i == 0 || std::isspace((unsigned char)s[i - 1])))
            return s.substr(0, i);
    }
    return s;
}

class FlowParser {
public:
    explicit FlowParser(const std::string& s) : s_(s) {}
    Node parse() {
        Node n = value();
        skip();
        if (p_ != s_.size()) fail("trailing characters");
        return n;
    }

private:
    const std::string& s_;
    size_t p_ = 0;
    [[noreturn]] void fail(const char* why) {
        throw std::runtime_error(std::string("flow collection: ") + why + " in '" + s_ + "'");
    }
    void skip() { while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_; }
    Node value() {
        skip();
        if (p_ < s_.size() && s_[p_] == '[') return seq();
        if (p_ < s_.size() && s_[p_] == '{') return mapping();
        Node n;
        size_t a = p_;
        int depth = 0;
        while (p_ < s_.size()) {
            char c = s_[p_];
            if (c == '(' ) ++depth;
            if (c == ')') --depth;
            if (depth <= 0 && (c == ',' || c == ']' || c == '}')) break;
            ++p_;
        }
        n.text = unquote(s_.substr(a, p_ - a));
        return n;
    }
    Node seq() {
        Node n;
        n.type = Node::SEQ;
        ++p_;
        skip();
        if (p_ < s_.size() && s_[p_] == ']') { ++p_; return n; }
        for (;;) {
            n.seq.push_back(value());
            skip();
            if (p_ >= s_.size()) fail("unterminated [");
            if (s_[p_] == ',') { ++p_; continue; }
            if (s_[p_] == ']') { ++p_; return n; }
            fail("expected , or ]");
        }
    }
    Node mapping() {
        Node n;
        n.type = Node::MAP;
        ++p_;
        skip();
        if (p_ < s_.size() && s_[p_] == '}') { ++p_; return n; }
        for (;;) {
            skip();
            size_t a = p_;
            while (p_ < s_.size() && s_[p_] != ':') ++p_;
            if (p_ >= s_.size()) fail("expected :");
            std::string key = unquote(s_.substr(a, p_ - a));
            ++p_;
            n.map.emplace_back(key, value());
            skip();
            if (p_ >= s_.size()) fail("unterminated {");
            if (s_[p_] == ',') { ++p_; continue; }
            if (s_[p_] == '}') { ++p_; return n; }
            fail("expected , or }");
        }
    }
};

struct Line {
    int indent;
    std::string text;   // without indentation
    int number;
};

class BlockParser {
public:
    BlockParser(std::vector<Line> lines, std::string file) : L_(std::move(lines)), file_(std::move(file)) {}
    Node parse() { return block(0); }

private:
    std::vector<Line> L_;
    std::string file_;
    size_t i_ = 0;

    [[noreturn]] void fail(int line, const std::string& why) {
        throw std::runtime_error(file_ + ":" + std::to_string(line) + ": " + why);
    }

    // Parse the block whose first line is at L_[i_] (indent >= min_indent).
    Node block(int min_indent) {
        Node n;
        if (i_ >= L_.size() || L_[i_].indent < min_indent) return n;
        int ind = L_[i_].indent;
        if (L_[i_].text.rfind("- ", 0) == 0 || L_[i_].text == "-") {
            n.type = Node::SEQ;
            while (i_ < L_.size() && L_[i_].indent == ind &&
                   (L_[i_].text.rfind("- ", 0) == 0 || L_[i_].text == "-")) {
                // Re-read the item body as a line indented two further.
                Line item = L_[i_];
                std::string rest = item.text.size() > 2 ? item.text.substr(2) : "";
                size_t lead = 0;
                while (lead < rest.size() && rest[lead] == ' ') ++lead;
                rest = rest.substr(lead);
                if (rest.empty()) {
                    ++i_;
                    n.seq.push_back(block(ind + 1));
                } else {
                    L_[i_].indent = ind + 2 + (int)lead;
                    L_[i_].text = rest;
                    n.seq.push_back(block(L_[i_].indent));
                }
            }
            return n;
        }
        // mapping or lone scalar
        if (!key_of(L_[i_].text)) {
            n = parse_inline(L_[i_].text, L_[i_].number);
            ++i_;
            return n;
        }
        n.type = Node::MAP;
        while (i_ < L_.size() && L_[i_].indent == ind) {
            const Line& ln = L_[i_];
            std::string key, rest;
            if (!split_key(ln.text, key, rest)) fail(ln.number, "expected 'key: value'");
            ++i_;
            if (rest.empty()) {
                if (i_ < L_.size() && (L_[i_].indent > ind ||
                    (L_[i_].indent == ind && L_[i_].text.rfind("- ", 0) == 0)))
                    n.map.emplace_back(key, block(L_[i_].indent));
                else
                    n.map.emplace_back(key, Node{});
            } else {
                n.map.emplace_back(key, parse_inline(rest, ln.number));
            }
        }
        if (i_ < L_.size() && L_[i_].indent > ind) fail(L_[i_].number, "unexpected indentation");
        return n;
    }

    static bool key_of(const std::string& t) {
        std::string k, r;
        return split_key(t, k, r);
    }
    static bool split_key(const std::string& t, std::string& key, std::string& rest) {
        if (t.empty() || t[0] == '[' || t[0] == '{') return false;
        bool sq = false, dq = false;
        for (size_t i = 0; i < t.size(); ++i) {
            char c = t[i];
            if (c == '\'' && !dq) sq = !sq;
            else if (c == '"' && !sq) dq = !dq;
            else if (c == ':' && !sq && !dq && (i + 1 == t.size() || t[i + 1] == ' ')) {
                key = unquote(t.substr(0, i));
                rest = trim(t.substr(i + 1));
                return true;
            }
        }
        return false;
    }
    Node parse_inline(const std::string& v, int line) {
        try {
            if (!v.empty() && (v[0] == '[' || v[0] == '{')) return FlowParser(v).parse();
        } catch (const std::exception& e) {
            fail(line, e.what());
        }
        Node n;
        n.text = unquote(v);
        return n;
    }
};

Node parse_yaml_file(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open mechanism file '" + path + "'");
    std::vector<std::string> raw;
    std::string s;
    while (std::getline(in, s)) {
        if (!s.empty() && s.back() == '\r') s.pop_back();
        raw.push_back(s);
    }
    std::vector<Line> lines;
    for (size_t i = 0; i < raw.size(); ++i) {
        std::string t = strip_comment(raw[i]);
        size_t ind = 0;
        while (ind < t.size() && t[ind] == ' ') ++ind;
        std::string body = trim(t);
        if (body.empty()) continue;
        // Block scalar: skip every following line indented deeper than this key.
        if (body.size() >= 2 && (body.substr(body.size() - 1) == "|" || body.substr(body.size() - 2) == "|-" ||
                                 body.substr(body.size() - 1) == ">" || body.substr(body.size() - 2) == ">-")) {
            size_t colon = body.rfind(": ");
            if (colon != std::string::npos) {
                lines.push_back({(int)ind, body.substr(0, colon + 1), (int)i + 1});
                // list item "- note: |" keeps its item indentation for the children
                int base = (int)ind + (body.rfind("- ", 0) == 0 ? 2 : 0);
                while (i + 1 < raw.size()) {
                    const std::string& nx = raw[i + 1];
                    size_t k = 0;
                    while (k < nx.size() && nx[k] == ' ') ++k;
                    if (trim(nx).empty() || (int)k > base) ++i;
                    else break;
                }
                continue;
            }
        }
        // Flow collections spanning several lines are joined.
        int depth = 0;
        for (char c : body) depth += (c == '[' || c == '{') - (c == ']' || c == '}');
        int start = (int)i + 1;
        while (depth > 0 && i + 1 < raw.size()) {
            std::string more = trim(strip_comment(raw[++i]));
            for (char c : more) depth += (c == '[' || c == '{') - (c == ']' || c == '}');
            body += " " + more;
        }
        if (depth != 0) throw std::runtime_error(path + ":" + std::to_string(start) + ": unbalanced brackets");
        lines.push_back({(int)ind, body, start});
    }
    return BlockParser(std::move(lines), path).parse();
}

double to_double(const Node* n, const std::string& what) {
    if (!n || n->type != Node::SCALAR) throw std::runtime_error("missing number: " + what);
    std::istringstream is(n->text);
    double v;
    if (!(is >> v)) throw std::runtime_error("bad number '" + n->text + "' for " + what);
    return v;
}

// Value with an optional unit suffix ("1.6599e+04 cal/mol").
double energy_to_K(const Node* n, const std::string& default_unit, const std::string& what) {
    if (!n || n->type != Node::SCALAR) throw std::runtime_error("missing activation energy: " + what);
    std::istringstream is(n->text);
    double v;
    if (!(is >> v)) throw std::runtime_error("bad activation energy '" + n->text + "' for " + what);
    std::string unit;
    std::getline(is, unit);
    unit = trim(unit);
    if (unit.empty()) unit = default_unit;
    if (unit == "cal/mol") return v * kCalToJ * 1000.0 / kGasConstant;
    if (unit == "kcal/mol") return v * kCalToJ * 1.0e6 / kGasConstant;
    if (unit == "J/mol") return v * 1000.0 / kGasConstant;
    if (unit == "kJ/mol") return v * 1.0e6 / kGasConstant;
    if (unit == "J/kmol") return v / kGasConstant;
    if (unit == "K") return v;
    throw std::runtime_error("unsupported activation-energy unit '" + unit + "' for " + what);
}

void parse_side(const std::string& side, const Mechanism& m, std::vector<int>& out, bool& third_body,
                bool& falloff, const std::string& eq) {
    std::string s = side;
    // "(+M)" / "(+ M)" marks falloff.
    for (;;) {
        // Only "(+" opens a collider group; species names such as CH2(S)
        // carry parentheses of their own.
        size_t a = s.find("(+");
        if (a == std::string::npos) break;
        size_t b = s.find(')', a);
        if (b == std::string::npos) throw std::runtime_error("bad equation '" + eq + "'");
        std::string inner = trim(s.substr(a + 1, b - a - 1));
        std::string sp;
        for (char c : inner) if (c != '+' && c != ' ') sp += c;
        if (sp != "M") throw std::runtime_error("unsupported falloff collider in '" + eq + "'");
        falloff = true;
        s = s.substr(0, a) + s.substr(b + 1);
    }
    // Split on " + " (species names such as "CH2(S)" never contain " + ").
    std::vector<std::string> terms;
    size_t pos = 0;
    for (;;) {
        size_t k = s.find(" + ", pos);
        terms.push_back(trim(s.substr(pos, k == std::string::npos ? std::string::npos : k - pos)));
        if (k == std::string::npos) break;
        pos = k + 3;
    }
    for (auto& t : terms) {
        if (t.empty()) continue;
        if (t == "M") { third_body = true; continue; }
        int coef = 1;
        std::string name = t;
        size_t sp = t.find(' ');
        if (sp != std::string::npos) {
            std::string c = t.substr(0, sp);
            char* end = nullptr;
            double v = std::strtod(c.c_str(), &end);
            if (end && *end == 0) {
                if (v != std::floor(v) || v < 1)
                    throw std::runtime_error("non-integer stoichiometry in '" + eq + "'");
                coef = (int)v;
                name = trim(t.substr(sp + 1));
            }
        }
        int k = m.index(name);
        if (k < 0) throw std::runtime_error("unknown species '" + name + "' in '" + eq + "'");
        for (int c = 0; c < coef; ++c) out.push_back(k);
    }
}

}  // namespace

double atomic_weight(const std::string& e) {
    // IUPAC conventional standard atomic weights.
    static const std::map<std::string, double> w = {
        {"H", 1.008}, {"C", 12.011}, {"N", 14.007}, {"O", 15.999}, {"Ar", 39.95}, {"AR", 39.95}, {"He", 4.002602}, {"HE", 4.002602}};
    auto it = w.find(e);
    if (it == w.end()) throw std::runtime_error("no atomic weight for element '" + e + "'");
    return it->second;
}

int Mechanism::index(const std::string& name) const {
    for (size_t k = 0; k < species.size(); ++k)
        if (species[k] == name) return (int)k;
    return -1;
}

std::string Mechanism::topology() const {
    std::ostringstream o;
    o << "species:";
    for (auto& s : species) o << s << ',';
    o << ";reactions:" << reactions.size() << ';';
    for (auto& r : reactions) {
        o << r.kind << (r.reversible ? 'r' : 'i') << (r.troe ? r.troe : 0) << ':';
        for (int k : r.reactants) o << k << ',';
        o << '>';
        for (int k : r.products) o << k << ',';
        o << '|';
        for (auto& e : r.efficiencies) o << e.first << ',';
        o << ';';
    }
    return o.str();
}

Mechanism load_mechanism(const std::string& path) {
    Node root = parse_yaml_file(path);
    Mechanism m;
    m.path = path;
    if (root.type != Node::MAP) throw std::runtime_error(path + ": top level is not a mapping");
    // Units: Cantera's YAML default is m / kmol / J-per-kmol.
    double length_to_m = 1.0, quantity_to_kmol = 1.0;
    std::string ea_unit = "J/kmol";
    if (const Node* u = root.get("units")) {
        if (const Node* l = u->get("length")) {
            if (l->text == "cm") length_to_m = 0.01;
            else if (l->text == "m") length_to_m = 1.0;
            else if (l->text == "mm") length_to_m = 0.001;
            else throw std::runtime_error(path + ": unsupported length unit " + l->text);
        }
        if (const Node* q = u->get("quantity")) {
            if (q->text == "mol") quantity_to_kmol = 1e-3;
            else if (q->text == "kmol") quantity_to_kmol = 1.0;
            else throw std::runtime_error(path + ": unsupported quantity unit " + q->text);
        }
        if (const Node* e = u->get("activation-energy")) ea_unit = e->text;
        if (const Node* t = u->get("time"))
            if (t->text != "s") throw std::runtime_error(path + ": unsupported time unit " + t->text);
    }
    const Node* phases = root.get("phases");
    if (!phases || phases->type != Node::SEQ || phases->seq.empty())
        throw std::runtime_error(path + ": no phases");
    const Node& ph = phases->seq[0];
    const Node* el = ph.get("elements");
    const Node* sp = ph.get("species");
    if (!el || el->type != Node::SEQ || !sp || sp->type != Node::SEQ)
        throw std::runtime_error(path + ": phase needs elements and species lists");
    for (auto& e : el->seq) m.elements.push_back(e.text);
    for (auto& s : sp->seq) m.species.push_back(s.text);
    const int ns = m.ns();
    m.W.assign(ns, 0.0);
    m.composition.assign(ns, std::vector<double>(m.elements.size(), 0.0));
    m.Tmin.assign(ns, 0);
    m.Tmid.assign(ns, 0);
    m.Tmax.assign(ns, 0);
    m.low.resize(ns);
    m.high.resize(ns);
    std::vector<bool> seen(ns, false);
    const Node* species = root.get("species");
    if (!species || species->type != Node::SEQ) throw std::runtime_error(path + ": no species section");
    for (const Node& s : species->seq) {
        const Node* nm = s.get("name");
        if (!nm) throw std::runtime_error(path + ": species without name");
        int k = m.index(nm->text);
        if (k < 0) continue;  // defined but not in the phase
        seen[k] = true;
        const Node* comp = s.get("composition");
        if (!comp || comp->type != Node::MAP) throw std::runtime_error(path + ": species " + nm->text + " lacks composition");
        double w = 0;
        for (auto& kv : comp->map) {
            int e = -1;
            for (size_t j = 0; j < m.elements.size(); ++j) if (m.elements[j] == kv.first) e = (int)j;
            if (e < 0) throw std::runtime_error(path + ": species " + nm->text + " uses element " + kv.first + " not in phase");
            double n = to_double(&kv.second, "composition");
            m.composition[k][e] = n;
            w += n * atomic_weight(kv.first);
        }
        m.W[k] = w;
        const Node* th = s.get("thermo");
        if (!th) throw std::runtime_error(path + ": species " + nm->text + " lacks thermo");
        const Node* model = th->get("model");
        if (!model || model->text != "NASA7") throw std::runtime_error(path + ": species " + nm->text + " thermo is not NASA7");
        const Node* tr = th->get("temperature-ranges");
        const Node* data = th->get("data");
        if (!tr || tr->type != Node::SEQ || !data || data->type != Node::SEQ)
            throw std::runtime_error(path + ": species " + nm->text + " NASA7 incomplete");
        auto coeffs = [&](const Node& d) {
            if (d.type != Node::SEQ || d.seq.size() != 7)
                throw std::runtime_error(path + ": species " + nm->text + " NASA7 needs 7 coefficients");
            std::array<double, 7> a;
            for (int i = 0; i < 7; ++i) a[i] = to_double(&d.seq[i], "NASA7 coefficient");
            return a;
        };
        if (tr->seq.size() == 3 && data->seq.size() == 2) {
            m.Tmin[k] = to_double(&tr->seq[0], "T");
            m.Tmid[k] = to_double(&tr->seq[1], "T");
            m.Tmax[k] = to_double(&tr->seq[2], "T");
            m.low[k] = coeffs(data->seq[0]);
            m.high[k] = coeffs(data->seq[1]);
        } else if (tr->seq.size() == 2 && data->seq.size() == 1) {
            m.Tmin[k] = to_double(&tr->seq[0], "T");
            m.Tmax[k] = to_double(&tr->seq[1], "T");
            m.Tmid[k] = m.Tmax[k];
            m.low[k] = m.high[k] = coeffs(data->seq[0]);
        } else {
            throw std::runtime_error(path + ": species " + nm->text + " has unsupported NASA7 ranges");
        }
    }
    for (int k = 0; k < ns; ++k)
        if (!seen[k]) throw std::runtime_error(path + ": species " + m.species[k] + " has no definition");

    const Node* rx = root.get("reactions");
    if (rx && rx->type == Node::SEQ) {
        for (const Node& r : rx->seq) {
            Reaction R;
            const Node* eqn = r.get("equation");
            if (!eqn) throw std::runtime_error(path + ": reaction without equation");
            R.equation = eqn->text;
            std::string lhs, rhs;
            size_t a;
            if ((a = R.equation.find("<=>")) != std::string::npos) {
                lhs = R.equation.substr(0, a); rhs = R.equation.substr(a + 3);
            } else if ((a = R.equation.find("=>")) != std::string::npos) {
                lhs = R.equation.substr(0, a); rhs = R.equation.substr(a + 2); R.reversible = false;
            } else if ((a = R.equation.find(" = ")) != std::string::npos) {
                lhs = R.equation.substr(0, a); rhs = R.equation.substr(a + 3);
            } else {
                throw std::runtime_error(path + ": cannot parse equation '" + R.equation + "'");
            }
            bool tbl = false, tbr = false, fol = false, forr = false;
            parse_side(lhs, m, R.reactants, tbl, fol, R.equation);
            parse_side(rhs, m, R.products, tbr, forr, R.equation);
            const Node* type = r.get("type");
            std::string t = type ? type->text : "elementary";
            if (const Node* d = r.get("duplicate")) R.duplicate = d->text == "true";
            if (r.get("orders") || r.get("negative-A") || r.get("negative-orders"))
                throw std::runtime_error(path + ": unsupported reaction option in '" + R.equation + "'");
            auto arrhenius = [&](const Node* rc, int order, double& A, double& b, double& EaR) {
                if (!rc || rc->type != Node::MAP) throw std::runtime_error(path + ": '" + R.equation + "' lacks a rate constant");
                A = to_double(rc->get("A"), "A of " + R.equation);
                b = to_double(rc->get("b"), "b of " + R.equation);
                EaR = energy_to_K(rc->get("Ea"), ea_unit, R.equation);
                // A carries (concentration)^(1-order) / s.
                double conc = quantity_to_kmol / (length_to_m * length_to_m * length_to_m);
                A *= std::pow(conc, 1 - order);
            };
            const int order = (int)R.reactants.size();
            if (t == "elementary") {
                if (tbl || tbr || fol || forr) throw std::runtime_error(path + ": '" + R.equation + "' has M but no type");
                R.kind = Reaction::ELEMENTARY;
                arrhenius(r.get("rate-constant"), order, R.A, R.b, R.EaR);
            } else if (t == "three-body") {
                if (!tbl || !tbr) throw std::runtime_error(path + ": three-body '" + R.equation + "' lacks M");
                R.kind = Reaction::THREE_BODY;
                arrhenius(r.get("rate-constant"), order + 1, R.A, R.b, R.EaR);
            } else if (t == "falloff") {
                if (!fol || !forr) throw std::runtime_error(path + ": falloff '" + R.equation + "' lacks (+M)");
                R.kind = Reaction::FALLOFF;
                arrhenius(r.get("high-P-rate-constant"), order, R.A, R.b, R.EaR);
                arrhenius(r.get("low-P-rate-constant"), order + 1, R.A0, R.b0, R.Ea0R);
                if (const Node* tr = r.get("Troe")) {
                    R.troeA = to_double(tr->get("A"), "Troe A");
                    R.troeT3 = to_double(tr->get("T3"), "Troe T3");
                    R.troeT1 = to_double(tr->get("T1"), "Troe T1");
                    R.troe = 3;
                    if (tr->get("T2")) { R.troeT2 = to_double(tr->get("T2"), "Troe T2"); R.troe = 4; }
                }
            } else {
                throw std::runtime_error(path + ": unsupported reaction type '" + t + "' in '" + R.equation + "'");
            }
            if (const Node* ef = r.get("efficiencies")) {
                for (auto& kv : ef->map) {
                    int k = m.index(kv.first);
                    if (k < 0) continue;   // Cantera skips efficiencies of absent species
                    R.efficiencies.emplace_back(k, to_double(&kv.second, "efficiency"));
                }
            }
            if (const Node* de = r.get("default-efficiency")) R.default_efficiency = to_double(de, "default-efficiency");
            m.reactions.push_back(std::move(R));
        }
    }
    return m;
}

}  // namespace hf
