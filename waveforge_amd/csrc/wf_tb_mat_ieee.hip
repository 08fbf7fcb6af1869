// Temporally blocked kernels, instantiated for materials=true, fast division=false.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_mat_ieee, true, false)

WF_TB_DECLARE(wf_tb_launch_mat_ieee_hist) {
    return a.tile_i == 48 ? wf_tb::launch_depth_hist<48, true, false>(a, st)
                          : wf_tb::launch_depth_hist<32, true, false>(a, st);
}
