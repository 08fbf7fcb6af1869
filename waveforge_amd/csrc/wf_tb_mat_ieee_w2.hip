// Temporally blocked kernels, instantiated for materials=true, IEEE division, two columns per lane.
#include "wf_tb_kernel.hpp"

WF_TB_DEFINE(wf_tb_launch_mat_ieee_w2, true, wf::kDivIeee, 2)
