// Temporally blocked kernels, instantiated for materials=true, the verified fast quotient, one column per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_mat_fast, true, wf::kDivFast, 1)
