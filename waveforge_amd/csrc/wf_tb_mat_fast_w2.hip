// Temporally blocked kernels, instantiated for materials=true, fast division=true, two columns per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_mat_fast_w2, true, true, 2)
