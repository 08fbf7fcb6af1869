// Temporally blocked kernels, instantiated for materials=false, reciprocal products, one column per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_free_rcp, false, wf::kDivRcp, 1)
