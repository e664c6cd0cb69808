// Spread/interp kernels instantiated for double.
#define JFA_REAL double
#include "spreadinterp_impl.inc"
