// Drop-in for <boost/numeric/itl/itl.hpp> on the sparse solve path: itl::cg,
// itl::bicgstab, the iteration classes and itl::pc::{identity, diagonal, ic_0, ilu_0}.
#pragma once
#include "../../../mtl4gpu.hpp"
