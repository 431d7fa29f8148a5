// Drop-in for <boost/numeric/mtl/matrix/sparse_banded.hpp>: the type lives in mtl4gpu.hpp.
#pragma once
#include "../../../../mtl4gpu.hpp"
