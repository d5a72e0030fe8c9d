// double N=6 instance of the collision kernels (see collision_kernels.h).
#include "collision_kernels.h"
namespace dgfs {
DGFS_COLLIDE_INST(double, 6)
}
