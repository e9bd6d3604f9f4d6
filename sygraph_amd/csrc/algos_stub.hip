/* Entry points whose kernels are not built yet: they fail loudly. */
#include "internal.hpp"
extern "C" {
int sg_bc(sg_graph_t, int32_t, const int32_t*, float*) { sg::set_error("sg_bc: not built yet"); return SG_ERR_UNSUPPORTED; }
int sg_bfs_multi(sg_graph_t, int32_t, const int32_t*, int32_t*) { sg::set_error("sg_bfs_multi: not built yet"); return SG_ERR_UNSUPPORTED; }
}
