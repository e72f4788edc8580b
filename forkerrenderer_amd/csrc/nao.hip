// nao.hip -- neural ambient occlusion (placeholder translation unit).
#include "fgl_kernels.hpp"
