"""decontamination_amd: MI355X-native (gfx950) batch self-organising-map hot path.

Drop-in GPU implementation of the batch SOM path of lpsc-euclid/decontamination
(``SOM_Batch`` training and the ``SOM_Abstract`` BMU / activation-map / errors inference) behind
the reference's class API. The compute runs in hand-written HIP kernels (libsomhip.so, C ABI in
include/somhip.h) called through ctypes.
"""
from .som_abstract import SOM_Abstract, dataset_to_generator_builder
from .som_batch import SOM_Batch
from .runtime import Context, DeviceArray, Model, Accumulator, get_context, gpu_available
from ._lib import SomHipError

__version__ = '0.1.0'

__all__ = [
    'SOM_Abstract',
    'SOM_Batch',
    'dataset_to_generator_builder',
    'Context',
    'DeviceArray',
    'Model',
    'Accumulator',
    'get_context',
    'gpu_available',
    'SomHipError',
]
