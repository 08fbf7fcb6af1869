// Temporally blocked kernels, instantiated for materials=false, IEEE division, two columns per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_free_ieee_w2, false, wf::kDivIeee, 2)
