// Umbrella header of the GemmNode host layer.
#pragma once
#include "momms/composables.hpp"
#include "momms/hip_gemm.hpp"
#include "momms/hip_tree.hpp"
#include "momms/matrix.hpp"
#include "momms/thread_comm.hpp"
#include "momms/util.hpp"
