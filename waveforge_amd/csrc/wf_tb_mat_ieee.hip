// Temporally blocked kernels, instantiated for materials=true, IEEE division, one column per lane
// (and the state-history variant the adjoint's recompute uses).
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_mat_ieee, true, wf::kDivIeee, 1)

WF_TB_DECLARE(wf_tb_launch_mat_ieee_hist) {
    return a.tile_i == 48 ? wf_tb::launch_depth<48, true, wf::kDivIeee, true>(a, st)
                          : wf_tb::launch_depth<32, true, wf::kDivIeee, true>(a, st);
}
