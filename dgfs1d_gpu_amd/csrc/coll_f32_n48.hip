// float N=48 instance of the collision kernels (see collision_kernels.h).
#include "collision_kernels.h"
namespace dgfs {
DGFS_COLLIDE_INST(float, 48)
}
