// Spread/interp kernels instantiated for float.
#define JFA_REAL float
#include "spreadinterp_impl.inc"
