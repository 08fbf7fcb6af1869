// Temporally blocked kernels, instantiated for materials=true, reciprocal products, two columns per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_mat_rcp_w2, true, wf::kDivRcp, 2)
