"""decontamination_amd: MI355X-native (gfx950) batch self-organising-map hot path.

Drop-in GPU implementation of the batch SOM path of lpsc-euclid/decontamination
(``SOM_Batch`` training and the ``SOM_Abstract`` BMU / activation-map / errors inference) behind
the reference's class API, with the steps on either side of it: the ``SOM_PCA`` initialisation and
its ``Covariance`` pass, and the galaxy-number-density products of ``Decontamination_SOM``. The
compute runs in hand-written HIP kernels (libsomhip.so, C ABI in include/somhip.h) called through
ctypes.
"""
from .som_abstract import SOM_Abstract, dataset_to_generator_builder
from .som_batch import SOM_Batch
from .som_pca import SOM_PCA
from .covariance import Covariance
from .decontamination_som import Decontamination_SOM
from .runtime import Context, DeviceArray, Model, Accumulator, get_context, gpu_available
from ._lib import SomHipError

__version__ = '0.1.0'

__all__ = [
    'SOM_Abstract',
    'SOM_Batch',
    'SOM_PCA',
    'Covariance',
    'Decontamination_SOM',
    'dataset_to_generator_builder',
    'Context',
    'DeviceArray',
    'Model',
    'Accumulator',
    'get_context',
    'gpu_available',
    'SomHipError',
]
