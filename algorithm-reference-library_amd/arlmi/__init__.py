"""arlmi: ARL's Fourier-processing hot path (gridding, degridding, w-stacking, shifted FFT,
grid correction) on AMD MI355X, behind ARL's own function signatures.

Compute runs in libarlmi.so (hand-written HIP for gfx950), called through ctypes with
PyTorch-ROCm tensors holding device memory.
"""

from . import _lib
from .data_models import (WCS, PolarisationFrame, Image, GridData, ConvolutionFunction, Visibility, BlockVisibility,
                          GainTable, Skycomponent, SkyCoord)
from .kernels import (create_image, create_convolutionfunction_from_image, create_pswf_convolutionfunction,
                      convolutionfunction_from_arrays,
                      create_box_convolutionfunction, create_awterm_convolutionfunction,
                      apply_bounding_box_convolutionfunction, calculate_bounding_box_convolutionfunction)
from .gridding import (convolution_mapping, grid_visibility_to_griddata, grid_visibility_to_griddata_fast,
                       degrid_visibility_from_griddata, fft_griddata_to_image, fft_image_to_griddata,
                       create_griddata_from_image)
from .imaging import (invert_2d, predict_2d, invert_wstack_single, predict_wstack_single, normalize_sumwt,
                      shift_vis_to_image, image_phasecentre, predict_skycomponent_visibility,
                      create_image_from_visibility, Imager)
from .workflows import (imaging_contexts, imaging_context, invert_list_serial_workflow,
                        predict_list_serial_workflow, sum_invert_results)
from .fft_support import fft, ifft
from .plan import Plan
from .streaming import StreamingImager
from .coalesce import (convert_blockvisibility_to_visibility, convert_blocks, decoalesce_visibility,
                       coalesce_visibility, average_in_blocks)
from .calibration import (create_gaintable_from_blockvisibility, apply_gaintable, gain_row_indices, apply_gains_rows,
                          apply_gains_rows_adjoint, solve_gaintable)
from .differentiable import PredictFunction, GainsFunction, predict_with_gains, visibility_loss
from .batch import Autotuner, BatchImager
from .phaserotate import skycoord_to_lmn, separation, phaserotate_visibility
from .deconvolution import hogbom, deconvolve_cube
from .weighting import (grid_weight_to_griddata, griddata_merge_weights, griddata_reweight, weight_visibility,
                        weight_list_serial_workflow)

__version__ = "0.1.0"
