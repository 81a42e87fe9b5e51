/* ORACLE -- test infrastructure only.  IMEX restatement (filled in below). */
#include "ec_oracle.h"
