// Temporally blocked kernels, instantiated for materials=true, reciprocal products, one column per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_mat_rcp, true, wf::kDivRcp, 1)
