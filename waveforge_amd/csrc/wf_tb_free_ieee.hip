// Temporally blocked kernels, instantiated for materials=false, fast division=false, one column per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_free_ieee, false, false, 1)
