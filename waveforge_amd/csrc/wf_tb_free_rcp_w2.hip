// Temporally blocked kernels, instantiated for materials=false, reciprocal products, two columns per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_free_rcp_w2, false, wf::kDivRcp, 2)
