"""sdepy_amd: sdepy's Euler path-integration hot path on AMD MI355X.

A drop-in for the preset processes of sdepy (``lognorm_process``,
``ornstein_uhlenbeck_process``, ``hull_white_process``,
``full_heston_process``, ``merton_jumpdiff_process``, ...), its memoryless
sources (``wiener_source``, ``poisson_source``, ``cpoisson_source``) and its
``process`` array type: called on a timeline, a preset returns a ``process``
exactly as in sdepy, but the step loop, the draws and the cross-path
reductions run in hand-written HIP kernels for gfx950 (libsdegpu.so, C ABI in
include/sdegpu.h).

Beyond sdepy's API: ``sdepy_amd.greeks`` (pathwise Greeks from reverse-mode
Euler adjoint kernels, with a ``torch.autograd.Function`` surface) and
``sdepy_amd.batched`` (many heterogeneous simulations in one persistent
launch).
"""
from . import infrastructure
from .infrastructure import (process, source, wiener_source, poisson_source,
                             cpoisson_source, odd_wiener_source,
                             even_poisson_source, even_cpoisson_source,
                             norm_rv, uniform_rv, exp_rv, double_exp_rv,
                             rvmap, path_moments, piecewise)
from .montecarlo import montecarlo
from .integration import (wiener_process, lognorm_process,
                          ornstein_uhlenbeck_process, hull_white_process,
                          hull_white_1factor_process,
                          cox_ingersoll_ross_process, full_heston_process,
                          heston_process, jumpdiff_process,
                          merton_jumpdiff_process, kou_jumpdiff_process)

# sdepy's short names (shortcuts.py:71-120), without the kfunc wrapping
dw = wiener_source
dn = poisson_source
dj = cpoisson_source
odd_dw = odd_wiener_source
even_dn = even_poisson_source
even_dj = even_cpoisson_source
wiener = wiener_process
lognorm = lognorm_process
oruh = ornstein_uhlenbeck_process
hwff = hull_white_process
hw1f = hull_white_1factor_process
cir = cox_ingersoll_ross_process
heston_xy = full_heston_process
heston = heston_process
jumpdiff = jumpdiff_process
mjd = merton_jumpdiff_process
kou = kou_jumpdiff_process

__version__ = '0.1.0'
