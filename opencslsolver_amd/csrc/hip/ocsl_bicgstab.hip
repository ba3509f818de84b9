// ILU(0)-BiCGSTAB drivers (filled in below).
#include "ocsl_handle.h"
#include "ocsl_kernels.h"

namespace ocsl {

void ilu_decompose_device(ocsl_solver*, int, bool) {
    throw Error{OCSL_ERR_ARG, "ILU not implemented yet"};
}
void bicgstab_reference(ocsl_solver*, ocsl_report*) {
    throw Error{OCSL_ERR_ARG, "BiCGSTAB not implemented yet"};
}
void bicgstab_tolerance(ocsl_solver*, bool, const ocsl_options&, ocsl_report*) {
    throw Error{OCSL_ERR_ARG, "BiCGSTAB not implemented yet"};
}

}  // namespace ocsl
