// Persistent VM instantiated for FP32 storage and leaf routines, BF16 products with FP32 accumulation.
#include "vm_run.h"

PP_VM_INSTANTIATE(mixed, float, MODE_FP32, MODE_BF16)
