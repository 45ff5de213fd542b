"""tbplas_amd: the TBPLaS TBPM hot path on AMD Instinct MI355X (gfx950).

Drop-in replacements for the TBPM path of TBPLaS:

    Sample.build_ham_csr / build_ham_dxy / rescale_ham / set_magnetic_field
    Config
    Solver.calc_corr_dos / calc_corr_ac_cond / calc_corr_dc_cond
    Analyzer.calc_dos / calc_ac_cond / calc_dc_cond

The compute runs in hand-written HIP kernels behind the C ABI of
``include/tbplas_amd.h`` (``tbplas_amd/_build/libtbplas_amd.so``); PyTorch is
used only to hold device buffers.
"""
from .config import Config, read_config
from .sample import (Sample, make_graphene_sample, graphene_hops, supercell_hops,
                     add_anderson_disorder, apply_uniaxial_strain,
                     random_vacancies, auto_tile_nm, ac_tile_nm)
from .solver import Solver
from .analysis import Analyzer, window_hanning, window_exp, window_exp_ten
from . import states, tbpm

__version__ = "0.1.0"

__all__ = ["Config", "read_config", "Sample", "make_graphene_sample",
           "graphene_hops", "add_anderson_disorder", "apply_uniaxial_strain",
           "random_vacancies", "auto_tile_nm", "ac_tile_nm", "Solver", "Analyzer", "window_hanning",
           "window_exp", "window_exp_ten", "states", "tbpm"]
