// iface_multi.hip -- the multi-instance interface pass: the interface code of multiphase.hip
// compiled once more with its grid, multiphase parameters and material tables in LDS (filled per
// block from an instance table) instead of constant memory, so that one launch solves the
// interface faces of several contexts of a batch.
#define M2C_IFACE_MULTI 1
#include "multiphase.hip"
