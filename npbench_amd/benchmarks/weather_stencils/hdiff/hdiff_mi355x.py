"""MI355X implementation of hdiff_numpy.hdiff (same signature).

``hdiff`` is the FP64 drop-in, one fused kernel (bitwise the NumPy result);
VARIANTS are the FP32/BF16 storage forms.
"""

import functools

from npbench_amd import stencils


def hdiff(in_field, out_field, coeff):
    stencils.hdiff_dropin(in_field, out_field, coeff, precision="fp64")


VARIANTS = {
    "fp32": functools.partial(stencils.hdiff_dropin, precision="fp32"),
    "bf16": functools.partial(stencils.hdiff_dropin, precision="bf16"),
}
