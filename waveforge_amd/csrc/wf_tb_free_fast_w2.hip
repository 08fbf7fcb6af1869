// Temporally blocked kernels, instantiated for materials=false, fast division=true, two columns per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_free_fast_w2, false, true, 2)
