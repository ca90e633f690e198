/* ORACLE (test infrastructure only): FP64 twin of the float restatement. */
#define ORC_REAL_IS_FLOAT 0
#include "oracle_strand.inc"
