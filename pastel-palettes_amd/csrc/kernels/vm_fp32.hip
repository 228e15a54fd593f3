// Persistent VM instantiated for FP32 storage and MFMA.
#include "vm_run.h"

PP_VM_INSTANTIATE(fp32, float, MODE_FP32, MODE_FP32)
