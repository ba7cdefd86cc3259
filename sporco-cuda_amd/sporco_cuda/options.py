"""Option defaults of the solvers: one table for the drop-in API and the
test oracle.

The reference's Python layer takes a complete sporco ``Options`` object
(or dict) and reads these keys (sporco_cuda/cbpdn.pyx:57-79, 179-193,
334-341); sporco, which owns the defaults, is not a dependency here, so the
table restates its documented ConvBPDN / ConvBPDNGradReg defaults.  This
module imports nothing else of the package, so the oracle can load it by
file path.
"""

import copy

_DEFAULTS = {
    'Verbose': False, 'MaxMainIter': 1000, 'AbsStopTol': 0.0,
    'RelStopTol': 1e-3, 'RelaxParam': 1.8, 'rho': None,
    'AutoRho': {'Enabled': True, 'Period': 1, 'AutoScaling': True,
                'Scaling': 1000.0, 'RsdlRatio': 1.2, 'RsdlTarget': None,
                'StdResiduals': False},
    'AuxVarObj': False, 'HighMemSolve': False, 'NonNegCoef': False,
    'NoBndryCross': False, 'L1Weight': 1.0,
}


def default_options(opt=None, grad=False):
    """Complete option dict: sporco's ConvBPDN (``grad=False``) or
    ConvBPDNGradReg (``grad=True``) defaults, updated from ``opt``
    (nested ``AutoRho`` entries are merged)."""
    d = copy.deepcopy(_DEFAULTS)
    if grad:
        d['GradWeight'] = 1.0
    if opt is not None:
        for k in opt:
            if k == 'AutoRho':
                d['AutoRho'].update(dict(opt['AutoRho']))
            else:
                d[k] = opt[k]
    return d
