// A C++ consumer of include/hyperflow.h, built against libhyperflow_hip.so
// the way the reference's driver would link it (g++, no HIP compiler):
//  1. the ChemistryBatch wrapper that replaces chemistry_batch.{hpp,cpp}
//     (HyperFlow_V3/src/chemistry_batch.hpp:8-17), as printed in the
//     integration guide, against a direct hf_chem_advance on the same cells;
//  2. one Euler stage over an AMReX-style box table whose boxes live in
//     separate device allocations, placed so that later boxes sit at lower
//     addresses (hf_flux_stage_boxes), against hf_flux_stage on the same
//     boxes packed into one array.
// Prints one line per check and exits nonzero on a mismatch. With --compile-only
// (or no visible GPU) it only checks that the library loads and the wrapper
// constructs its error path.
#include "hyperflow.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

using State = std::vector<double>;
using Field = std::vector<State>;

// ---- the integration guide's wrapper, unchanged
class ChemistryBatch {
public:
    explicit ChemistryBatch(const std::string& mechanism, int /*threads*/ = 0) {
        if (hf_create(mechanism.c_str(), nullptr, &ctx) != HF_OK)       // rtol 1e-8, atol 1e-15, 5000 steps
            throw std::runtime_error(hf_create_error());
        nv = hf_num_species(ctx) + 3;
    }
    ~ChemistryBatch() { hf_destroy(ctx); }
    void advance(Field& field, double dt, size_t count) {                // chemistry_batch.cpp:21-50
        if (count > field.size()) throw std::invalid_argument("chemistry count exceeds field size");
        soa.resize(nv * count);
        for (size_t i = 0; i < count; ++i)
            for (int n = 0; n < nv; ++n) soa[n * count + i] = field[i][n];
        hf_chem_stats st{};
        int64_t bad = -1;
        const int r = hf_chem_advance(ctx, soa.data(), (int64_t)count, (int64_t)count, &dt, 0,
                                      HF_CHEM_BATCHED, &st, &bad);
        if (r == HF_EDOMAIN) throw std::domain_error(hf_last_error(ctx));   // BlockAMR.cpp:258 branches on this
        if (r != HF_OK) throw std::runtime_error(hf_last_error(ctx));       // field untouched (all-or-nothing)
        for (size_t i = 0; i < count; ++i)
            for (int n = 0; n < nv; ++n) field[i][n] = soa[n * count + i];
        lastTeam = st.team;
    }
    int threads() const { return 1; }
    int lastTeamSize() const { return lastTeam; }
private:
    hf_ctx* ctx = nullptr;
    int nv = 0, lastTeam = 1;
    std::vector<double> soa;
};

namespace {

int failures = 0;

void report(const char* what, bool ok, const std::string& detail = "") {
    std::printf("%s %s%s%s\n", ok ? "PASS" : "FAIL", what, detail.empty() ? "" : ": ", detail.c_str());
    std::fflush(stdout);
    if (!ok) ++failures;
}

#define HIPCHECK(x)                                                                        \
    do {                                                                                   \
        if ((x) != hipSuccess) throw std::runtime_error(std::string("HIP call failed: ") + #x); \
    } while (0)

void check_chemistry(const std::string& mech) {
    hf_ctx* ctx = nullptr;
    if (hf_create(mech.c_str(), nullptr, &ctx) != HF_OK) throw std::runtime_error(hf_create_error());
    const int nv = hf_num_species(ctx) + 3;
    const int64_t n = 300;
    std::vector<double> U(nv * n), dts(n);
    if (hf_synth_cells(ctx, 0, 2, n, U.data(), dts.data()) != HF_OK) throw std::runtime_error(hf_last_error(ctx));
    const double dt = 1e-7;
    // direct call on the SoA planes
    std::vector<double> direct = U;
    hf_chem_stats st{};
    int64_t bad = -1;
    const int r = hf_chem_advance(ctx, direct.data(), n, n, &dt, 0, HF_CHEM_BATCHED, &st, &bad);
    report("hf_chem_advance on host planes", r == HF_OK, r == HF_OK ? "" : hf_last_error(ctx));
    hf_destroy(ctx);
    // the wrapper on the reference's AoS field
    Field field(n, State(nv));
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < nv; ++c) field[i][c] = U[c * n + i];
    ChemistryBatch batch(mech);
    batch.advance(field, dt, (size_t)n);
    bool same = true;
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < nv; ++c) same = same && field[i][c] == direct[c * n + i];
    report("ChemistryBatch wrapper equals the direct call bit for bit", same);
    // the wrapper's error contract: a count beyond the field and an inadmissible cell
    bool threw = false;
    try {
        batch.advance(field, dt, field.size() + 1);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    report("ChemistryBatch rejects a count beyond the field", threw);
    Field badf(2, State(nv, 0.0));
    threw = false;
    try {
        batch.advance(badf, dt, 2);
    } catch (const std::domain_error&) {
        threw = true;
    }
    report("ChemistryBatch throws domain_error for an inadmissible cell", threw);
}

// A smooth admissible field: blocks of 4 x 4 cells share a synthetic state,
// with a velocity field added (kinetic energy added to E).
void fill_box(const std::vector<double>& S, int ns_cells, int nv, int ns, int gx, int gy, int seed, double* out,
              int64_t pitch, int64_t plane) {
    for (int j = 0; j < gy; ++j)
        for (int i = 0; i < gx; ++i) {
            const int k = ((i / 4) + 7 * (j / 4) + seed) % ns_cells;
            double rho = 0;
            for (int s = 0; s < ns; ++s) rho += S[s * ns_cells + k];
            const double u = 200.0 * std::sin(0.3 * i + seed), v = 100.0 * std::cos(0.2 * j + seed);
            for (int c = 0; c < nv; ++c) out[c * plane + j * pitch + i] = S[c * ns_cells + k];
            out[ns * plane + j * pitch + i] = rho * u;
            out[(ns + 1) * plane + j * pitch + i] = rho * v;
            out[(ns + 2) * plane + j * pitch + i] += 0.5 * rho * (u * u + v * v);
        }
}

void check_box_table(const std::string& mech) {
    hf_ctx* ctx = nullptr;
    if (hf_create(mech.c_str(), nullptr, &ctx) != HF_OK) throw std::runtime_error(hf_create_error());
    const int ns = hf_num_species(ctx), nv = ns + 3;
    const int nsyn = 16;
    std::vector<double> S(nv * nsyn);
    if (hf_synth_cells(ctx, 0, 3, nsyn, S.data(), nullptr) != HF_OK) throw std::runtime_error(hf_last_error(ctx));
    const int sizes[3][2] = {{20, 12}, {9, 17}, {33, 5}};
    const int walls[3] = {0, 1 | 4, 15};
    // separate allocations: three ghosted inputs, three outputs
    struct Fab { double* in; double* out; int nx, ny; };
    std::vector<Fab> fabs;
    for (int b = 0; b < 3; ++b) {
        Fab f{nullptr, nullptr, sizes[b][0], sizes[b][1]};
        HIPCHECK(hipMalloc((void**)&f.in, sizeof(double) * nv * (f.nx + 4) * (f.ny + 4)));
        HIPCHECK(hipMalloc((void**)&f.out, sizeof(double) * nv * f.nx * f.ny));
        fabs.push_back(f);
    }
    // later boxes at lower addresses: the order a pointer difference from the
    // first box would make negative
    std::sort(fabs.begin(), fabs.end(), [](const Fab& a, const Fab& b) { return a.in > b.in; });
    std::vector<hf_box> boxes;
    std::vector<const double*> ins;
    std::vector<double*> outs;
    std::vector<std::vector<double>> host_in;
    int64_t packed_in = 0, packed_out = 0;
    for (size_t b = 0; b < fabs.size(); ++b) {
        const Fab& f = fabs[b];
        const int gx = f.nx + 4, gy = f.ny + 4;
        host_in.emplace_back((size_t)nv * gx * gy);
        fill_box(S, nsyn, nv, ns, gx, gy, (int)b, host_in.back().data(), gx, (int64_t)gx * gy);
        HIPCHECK(hipMemcpy(f.in, host_in.back().data(), sizeof(double) * nv * gx * gy, hipMemcpyHostToDevice));
        HIPCHECK(hipMemset(f.out, 0, sizeof(double) * nv * f.nx * f.ny));
        hf_box B{};
        B.nx = f.nx; B.ny = f.ny; B.walls = walls[b]; B.wedge_i = 0;
        B.in_offset = 0; B.in_pitch = gx; B.in_plane = (int64_t)gx * gy;
        B.out_offset = 0; B.out_pitch = f.nx; B.out_plane = (int64_t)f.nx * f.ny;
        B.dtdx = 2e-4; B.dtdy = 1.5e-4;                 // dt/dx, dt/dy in s/m
        boxes.push_back(B);
        ins.push_back(f.in);
        outs.push_back(f.out);
        packed_in += (int64_t)gx * gy;
        packed_out += (int64_t)f.nx * f.ny;
    }
    hf_flux_stats st{};
    int64_t bad = -1;
    int r = hf_flux_stage_boxes(ctx, boxes.data(), (int)boxes.size(), ins.data(), outs.data(), 0, nullptr, &st, &bad);
    report("hf_flux_stage_boxes over separate allocations in descending address order", r == HF_OK,
           r == HF_OK ? "" : hf_last_error(ctx));
    // the same boxes packed into one input and one output array (plane = total)
    std::vector<double> pin((size_t)nv * packed_in), pout((size_t)nv * packed_out, 0.0);
    std::vector<hf_box> pboxes = boxes;
    int64_t oi = 0, oo = 0;
    for (size_t b = 0; b < boxes.size(); ++b) {
        const int64_t cells = boxes[b].in_plane;
        for (int c = 0; c < nv; ++c)
            std::memcpy(&pin[c * packed_in + oi], &host_in[b][c * cells], sizeof(double) * cells);
        pboxes[b].in_offset = oi; pboxes[b].in_plane = packed_in;
        pboxes[b].out_offset = oo; pboxes[b].out_plane = packed_out;
        oi += cells;
        oo += boxes[b].out_plane;
    }
    double *din = nullptr, *dout = nullptr;
    HIPCHECK(hipMalloc((void**)&din, sizeof(double) * pin.size()));
    HIPCHECK(hipMalloc((void**)&dout, sizeof(double) * pout.size()));
    HIPCHECK(hipMemcpy(din, pin.data(), sizeof(double) * pin.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(dout, 0, sizeof(double) * pout.size()));
    r = hf_flux_stage(ctx, pboxes.data(), (int)pboxes.size(), din, dout, 0, &st, &bad);
    report("hf_flux_stage on the packed boxes", r == HF_OK, r == HF_OK ? "" : hf_last_error(ctx));
    HIPCHECK(hipMemcpy(pout.data(), dout, sizeof(double) * pout.size(), hipMemcpyDeviceToHost));
    bool same = true;
    oo = 0;
    for (size_t b = 0; b < boxes.size(); ++b) {
        std::vector<double> got((size_t)nv * boxes[b].out_plane);
        HIPCHECK(hipMemcpy(got.data(), outs[b], sizeof(double) * got.size(), hipMemcpyDeviceToHost));
        for (int c = 0; c < nv; ++c)
            for (int64_t k = 0; k < boxes[b].out_plane; ++k)
                same = same && got[c * boxes[b].out_plane + k] == pout[c * packed_out + oo + k];
        oo += boxes[b].out_plane;
    }
    report("per-box pointers give the packed stage's update bit for bit", same);
    // a box that reads past its own allocation is refused
    std::vector<hf_box> over = boxes;
    over[0].in_offset = 8;
    r = hf_flux_stage_boxes(ctx, over.data(), (int)over.size(), ins.data(), outs.data(), 0, nullptr, &st, &bad);
    report("a box reading past its allocation is refused with HF_EINVAL", r == HF_EINVAL);
    for (auto& f : fabs) {
        HIPCHECK(hipFree(f.in));
        HIPCHECK(hipFree(f.out));
    }
    HIPCHECK(hipFree(din));
    HIPCHECK(hipFree(dout));
    hf_destroy(ctx);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: abi_consumer MECHANISM.yaml [--compile-only]\n");
        return 2;
    }
    const std::string mech = argv[1];
    const bool gpu = !(argc > 2 && std::strcmp(argv[2], "--compile-only") == 0);
    try {
        if (!gpu) {
            // no device work: the loader and the wrapper's error path
            bool threw = false;
            try {
                ChemistryBatch missing(mech + ".missing");
            } catch (const std::runtime_error&) {
                threw = true;
            }
            report("ChemistryBatch on a missing mechanism throws runtime_error", threw);
            return failures ? 1 : 0;
        }
        check_chemistry(mech);
        check_box_table(mech);
    } catch (const std::exception& e) {
        report("exception", false, e.what());
    }
    return failures ? 1 : 0;
}
