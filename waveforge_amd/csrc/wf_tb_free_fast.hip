// Temporally blocked kernels, instantiated for materials=false, the verified fast quotient, one column per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_free_fast, false, wf::kDivFast, 1)
