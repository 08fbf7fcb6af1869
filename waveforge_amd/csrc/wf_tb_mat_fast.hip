// Temporally blocked kernels, instantiated for materials=true, fast division=true, one column per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_mat_fast, true, true, 1)
