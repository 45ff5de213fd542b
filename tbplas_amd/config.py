"""Drop-in ``Config`` for the TBPM path (tbplas/tbpm/config.py:11-185).

Same parameter groups, keys and defaults as the reference, the same
``set_temperature`` / ``check_params`` / ``read_config`` behaviour, plus a
``gpu`` group for the MI355X path that the reference's key check does not
look at:

    gpu['precision']   'fp64' (complex128) or 'fp32' (complex64 storage and
                       arithmetic, FP64 correlation sums)     default 'fp64'
    gpu['device']      GPU index                              default 0
    gpu['use_graph']   replay time steps from captured hipGraphs  default True
    gpu['autotune']    time the candidate launch shapes once per problem
                       shape and keep the fastest (tbplas_amd.tune) default True
    gpu['max_vectors'] states propagated together per launch  default 16
    gpu['state_kind']  'phase' (unit amplitude, random phase) or 'exp'
                       (the reference's Exp(1) amplitudes)    default 'phase'
    gpu['fp32_norm_thr'] norm-check threshold used on the fp32 path, which
                       drifts ~1e-6 per hundred steps         default 1e-4
    gpu['tile_nm']     DOS, LDOS and quasi-eigenstate runs: store the rows
                       in square tiles of this size (nm, from the orbital
                       positions) so the tuner may run three to five
                       Chebyshev terms per pass; same results, 0 keeps the
                       sample's order, 'auto' sizes the tiles for the
                       states per row (sample.auto_tile_nm)   default 'auto'
    gpu['cond_tile_nm'] the same for dynamic-polarisation runs (more
                       vectors per row, so smaller tiles)     default 1.5
    gpu['ac_states']   states per AC run (three propagated vectors each):
                       fewer vectors per row allow larger tiles
                       (8.4M-orbital graphene, complex64: 4 states on
                       2.5 nm tiles 638 state-steps/s, 8 on 1.5 nm 519)
                                                              default 4
    gpu['ac_tile_nm']  the tile edge of AC runs ('auto': sample.ac_tile_nm
                       for the run's vectors, 3.8 nm for 4 complex64
                       states on graphene, 2.5 nm for complex128)
                                                              default 'auto'
    gpu['dc_states']   states per DC run (two propagated vectors each)
                       (2M-orbital disordered graphene, complex64,
                       step-vectors/s: 8 states on 3.2 nm tiles 4696,
                       16 on 2.0 nm 4312, 4 on 4.8 nm 4632)   default 8
    gpu['dc_tile_nm']  the tile edge of DC runs ('auto': auto_tile_nm for
                       the run's vectors, 3.2 nm for 8 complex64 states
                       on graphene)                           default 'auto'
"""
import pickle

__all__ = ["Config", "read_config", "KB"]

KB = 8.617333262e-5  # Boltzmann constant in eV/K (base/constants.py:7)


class Config:
    """TBPM parameters (tbplas/tbpm/config.py:84-124)."""

    def __init__(self):
        self.generic = {'Bessel_max': 250,
                        'Bessel_precision': 1.0e-14,
                        'beta': 1.0 / (KB * 300),
                        'mu': 0.,
                        'correct_spin': False,
                        'nr_time_steps': 1024,
                        'nr_random_samples': 1,
                        'seed': 1337,
                        'wfn_check_steps': 128,
                        'wfn_check_thr': 1.0e-9,
                        'nr_Fermi_fft_steps': 2 ** 15,
                        'Fermi_cheb_precision': 1.0e-12}
        self.LDOS = {'site_indices': [0],
                     'wf_weights': [1.0],
                     'delta': 0.01,
                     'recursion_depth': 2000}
        self.DC_conductivity = {'energy_limits': (-0.5, 0.5)}
        self.quasi_eigenstates = {'energies': [-0.1, 0., 0.1]}
        self.dyn_pol = {'q_points': [[1., 0., 0.]],
                        'coulomb_constant': 1.0,
                        'background_dielectric_constant': 1.0}
        self.dckb = {'energies': [i * 0.01 - 0.2 for i in range(0, 41)],
                     'n_kernel': 2048,
                     'direction': 1,
                     'ne_integral': 2048}
        self.gpu = {'precision': 'fp64',
                    'device': 0,
                    'use_graph': True,
                    'autotune': True,
                    'max_vectors': 16,
                    'state_kind': 'phase',
                    'fp32_norm_thr': 1.0e-4,
                    'tile_nm': 'auto',
                    'cond_tile_nm': 1.5,
                    'ac_states': 4,
                    'ac_tile_nm': 'auto',
                    'dc_states': 8,
                    'dc_tile_nm': 'auto'}
        self._legal_params = dict()
        self.set_legal_params()

    def set_temperature(self, temperature=300.0):
        """beta = 1 / (kB T) (config.py:126-132)."""
        self.generic['beta'] = 1.0 / (KB * temperature)

    def set_legal_params(self):
        self._legal_params = {
            'generic': set(self.generic.keys()),
            'LDOS': set(self.LDOS.keys()),
            'dyn_pol': set(self.dyn_pol.keys()),
            'dckb': set(self.dckb.keys()),
        }

    def check_params(self):
        """Raise ValueError on unknown keys in the groups the reference checks
        (config.py:147-165), and on an unknown gpu precision."""
        for name in ('generic', 'LDOS', 'dyn_pol', 'dckb'):
            unknown = set(getattr(self, name).keys()) - self._legal_params[name]
            for key in unknown:
                raise ValueError(f"Undefined parameter {key} in config.{name}")
        gpu = getattr(self, 'gpu', {})
        if gpu.get('precision', 'fp64') not in ('fp64', 'fp32'):
            raise ValueError(f"Unknown precision {gpu.get('precision')!r}")


def read_config(filename):
    """Read a Config pickled by Solver.save_config (config.py:168-185).

    Only load files this package wrote: pickle executes code from the file.
    """
    with open(filename, 'rb') as f:
        saved = pickle.load(f)
    config = Config()
    for name in ('generic', 'LDOS', 'dyn_pol', 'DC_conductivity', 'dckb',
                 'quasi_eigenstates'):
        setattr(config, name, getattr(saved, name))
    if hasattr(saved, 'gpu'):
        config.gpu.update(saved.gpu)
    config.set_legal_params()
    return config
