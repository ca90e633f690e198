/* ORACLE (test infrastructure only): float expression runtime restatement
 * (src/xgen_expression.cpp).  Filled in with the expression path. */
#include "oracle.h"
