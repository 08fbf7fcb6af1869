// Temporally blocked kernels, instantiated for materials=false, IEEE division, one column per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_free_ieee, false, wf::kDivIeee, 1)
