// float N=48 instance of the binary-mixture VSS kernels (see bivss_kernels.h).
#include "bivss_kernels.h"
namespace dgfs {
DGFS_BI_INST(float, 48)
}
