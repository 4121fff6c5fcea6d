"""mpk: the MI355X Krylov core behind `mojopetsc4py.PETSc`.

`mpk._native` loads libmpk.so (hand-written gfx950 HIP kernels behind a C
ABI), `mpk.device` owns device memory, `mpk.workloads` builds the synthetic
systems the tests and the benchmark use.
"""

from ._native import (NativeError, check, device_name, lib, reduction_order,  # noqa: F401
                      reduction_order_scope, set_reduction_order)
