// Persistent VM instantiated for FP64 storage and MFMA.
#include "vm_run.h"

PP_VM_INSTANTIATE(fp64, double, MODE_FP64, MODE_FP64)
