"""GraphFlood warm-start surrogate: a small convolutional U-Net that predicts a
steady water-depth field from [DEM, log drainage area, rainfall, Manning n]
so that GraphFlood starts near its steady state.  No reference counterpart.

The network's convolutions, pooling, upsampling, concatenation and head run in
the HIP kernels of libpff_hip.so (BF16 activations, FP32 accumulation on the
matrix cores); torch provides tensors, autograd bookkeeping and the optimiser.
"""

from .ops import avgpool2, concat, conv3x3, head1x1, pack_input, upsample2  # noqa: F401
from .unet import UNet  # noqa: F401
