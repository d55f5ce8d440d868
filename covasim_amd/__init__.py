"""covasim_amd: Covasim's daily transmission step on AMD MI355X (gfx950).

A drop-in for the reference's transmission path behind a ``cv.Sim`` / ``cv.People``
shaped API: ``utils.compute_viral_load``, ``utils.compute_trans_sus``,
``utils.compute_infections``, ``People.infect`` and ``People.update_states_pre``
run as hand-written HIP kernels (libcovasim_amd.so, C ABI in include/covasim_amd.h).
"""

from .version import __version__
from .parameters import make_pars, reset_layer_pars, get_prognoses, get_variant_pars, get_cross_immunity
from .immunity import variant
from .base import Result, Layer, Contacts
from .people import People
from .sim import Sim, AlreadyRunError
from .batch import BatchSim
from .run import MultiSim, single_run, multi_run
from .interventions import Intervention, change_beta, clip_edges, test_prob, contact_tracing
from .checkpoint import load
from . import utils
from .utils import set_seed, choose, choose_r, default_int, default_float

__all__ = ['__version__', 'Intervention', 'change_beta', 'clip_edges', 'test_prob', 'contact_tracing', 'load', 'make_pars', 'reset_layer_pars', 'get_prognoses', 'get_variant_pars',
           'get_cross_immunity', 'variant', 'Result', 'Layer', 'Contacts', 'People', 'Sim', 'AlreadyRunError', 'BatchSim', 'MultiSim', 'single_run', 'multi_run',
           'utils', 'set_seed', 'choose', 'choose_r', 'default_int', 'default_float']
