// Umbrella header of the MI355X einsum backend's C++ API.
#pragma once

#include "DeviceTensor.hpp"
#include "DiskTensor.hpp"
#include "Index.hpp"
#include "Tensor.hpp"
#include "TensorAlgebra.hpp"
