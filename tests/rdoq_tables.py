"""Synthetic coefficient rate tables for the RDOQ tests: the generator of
the package's synth module (random per-symbol costs, secondary entries
derived as the reference's table builder does)."""
import importlib.util
import os

_p = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'zenav1-svt_amd', 'synth.py')
_spec = importlib.util.spec_from_file_location('zav1_synth_tables', _p)
_m = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_m)

TXB_WORDS, EOB_WORDS = _m.TXB_WORDS, _m.EOB_WORDS
SKIP, BASE_EOB, BASE, EOB_EXTRA, DC_SIGN, LPS = _m.SKIP, _m.BASE_EOB, _m.BASE, _m.EOB_EXTRA, _m.DC_SIGN, _m.LPS
coeff_costs = _m.coeff_costs
tables_for = _m.tables_for
