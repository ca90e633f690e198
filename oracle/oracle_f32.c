/* ORACLE (test infrastructure only): float restatement of generate.h. */
#define ORC_REAL_IS_FLOAT 1
#include "oracle_strand.inc"
