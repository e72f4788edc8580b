// graph.cpp -- (placeholder translation unit)
#include <hip/hip_runtime.h>
