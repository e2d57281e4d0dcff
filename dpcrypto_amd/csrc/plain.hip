// Plain one-workgroup-per-KEM comparator kernels (filled in below).
#include "common.hpp"
