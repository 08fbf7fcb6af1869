// Temporally blocked kernels, instantiated for materials=true, the verified fast quotient, two columns per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_mat_fast_w2, true, wf::kDivFast, 2)
