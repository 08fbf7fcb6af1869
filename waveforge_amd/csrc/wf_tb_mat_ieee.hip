// Temporally blocked kernels, instantiated for materials=true, fast division=false.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_mat_ieee, true, false)
