"""subby_amd: MI355X (gfx950) engine for subby's per-column sufficient statistics.

Public surface mirrors subby.jax for the single-model path:
LogLike, Counts, RootProb, BranchCounts, InsideOutside, LogLikeCustomGrad,
Tree, DiagModel, RateModel and the model builders.  Kernels live in libsubby_amd.so (built by
``python -m subby_amd.build``); there is no CPU fallback.
"""
from .models import (Tree, DiagModel, IrrevDiagModel, RateModel, hky85_diag, hky85_rate_matrix,
                     hky85_log_kappa_weights, jukes_cantor_model, f81_model,
                     diagonalize_rate_matrix, reversible_rate_matrix, ensure_diag,
                     diagonalize_irreversible, irrev_model_from_rate_matrix,
                     model_from_rate_matrix, check_detailed_balance, gamma_rate_categories,
                     scale_model, genetic_code, codon_to_sense, gy94_rate_matrix, gy94_model,
                     diagonalize_rate_matrix_auto)
from ._lib import SubbyError
from .api import (Engine, default_engine, LogLike, Counts, RootProb, loglike_grad, BranchCounts,
                  InsideOutside, MixturePosterior)
from .grad import LogLikeCustomGrad, hky85_loglike

__all__ = ['Tree', 'DiagModel', 'IrrevDiagModel', 'RateModel',
           'diagonalize_irreversible', 'irrev_model_from_rate_matrix', 'model_from_rate_matrix',
           'check_detailed_balance', 'gamma_rate_categories', 'scale_model', 'MixturePosterior',
           'genetic_code', 'codon_to_sense', 'gy94_rate_matrix', 'gy94_model',
           'pad_alignment', 'unpad_columns', 'pad_tree', 'pad_tree_and_alignment', 'pad_to_geom_bin', 'hky85_diag', 'hky85_rate_matrix',
           'hky85_log_kappa_weights', 'jukes_cantor_model', 'f81_model',
           'diagonalize_rate_matrix', 'reversible_rate_matrix', 'ensure_diag', 'SubbyError',
           'Engine', 'default_engine', 'LogLike', 'Counts', 'RootProb', 'loglike_grad',
           'BranchCounts', 'InsideOutside', 'GEOM_BINS', 'diagonalize_rate_matrix_auto',
           'LogLikeCustomGrad', 'hky85_loglike']
from .padding import (GEOM_BINS, pad_alignment, unpad_columns, pad_tree,  # noqa: E402
                      pad_tree_and_alignment, pad_to_geom_bin)
from . import synthetic, batch  # noqa: E402
