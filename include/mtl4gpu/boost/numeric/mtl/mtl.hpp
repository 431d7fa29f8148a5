// Drop-in for <boost/numeric/mtl/mtl.hpp> on the sparse solve path: the
// GPU-backed compressed2D / dense_vector and their operations.
#pragma once
#include "../../../mtl4gpu.hpp"
