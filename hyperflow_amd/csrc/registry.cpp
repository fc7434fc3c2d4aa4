// Table of the mechanism topologies compiled into this library. The build
// writes registry.inc with one HF_TOPOLOGY(name) line per mechanism.
#include "topo_ops.hpp"

namespace hf {
#define HF_TOPOLOGY(name) extern const TopoOps ops_##name;
#include "registry.inc"
#undef HF_TOPOLOGY

const TopoOps* const* compiled_topologies() {
#define HF_TOPOLOGY(name) &ops_##name,
    static const TopoOps* const table[] = {
#include "registry.inc"
        nullptr};
#undef HF_TOPOLOGY
    return table;
}
}  // namespace hf
