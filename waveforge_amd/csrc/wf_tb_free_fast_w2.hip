// Temporally blocked kernels, instantiated for materials=false, the verified fast quotient, two columns per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_free_fast_w2, false, wf::kDivFast, 2)
