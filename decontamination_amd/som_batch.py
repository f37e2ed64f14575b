"""Batch SOM with the reference's API, trained by the HIP kernels of libsomhip.

Mirrors ``decontamination.algo.som_batch.SOM_Batch`` (decontamination/algo/som_batch.py:29-505):
constructor ``(m, n, dim, dtype, topology, sigma)``, ``sigma`` property, and ``train`` with the same
arguments and results (``weights``/``centroids``, ``_history``, per-epoch quantization and
topographic errors, early stop, best epoch).

The epoch keeps the reference's algorithm (som_batch.py:217-228) in its factorised form: BMU search
on the GPU, exact per-BMU sums S_c = sum rho_i x_i and c_c = sum rho_i, then num = H.S and
den = H.c with H_jc = exp(-|t_j - t_c|^2 / (2 sigma^2)), and w = num/den where den > 1e-12 else 0.
This equals the reference's per-sample scatter because h depends only on (j, bmu(i)).

Errors of epoch e (computed by the reference after the update, som_batch.py:363) are the errors of
the weights W(e+1); they are fused into the BMU pass of epoch e+1, which runs on the same weights,
plus one final errors pass. Early stopping (stop_* given) uses an explicit errors pass per epoch
instead, as it needs the value before the next epoch starts.
"""
import math
import typing

import numpy as np

from . import _lib
from .runtime import Accumulator, DeviceArray
from .som_abstract import SOM_Abstract, iterate_chunks

FLOAT_EPS = 1.0e-12


# noinspection PyPep8Naming
class SOM_Batch(SOM_Abstract):
    """Self Organizing Maps (standard batch implementation), GPU.

    Parameters: m, n, dim, dtype (np.float32), topology (None means hexagonal), sigma (None means
    max(m, n)/2), plus the keyword-only options of ``SOM_Abstract`` (precision, arithmetic, device).
    """

    __MODE__ = 'batch'

    def __init__(self, m: int, n: int, dim: int, dtype=np.float32, topology: typing.Optional[str] = None,
                 sigma: float = None, **options):
        super().__init__(m, n, dim, dtype, topology, **options)
        self._sigma = max(m, n) / 2.0 if sigma is None else float(sigma)
        self._n_epochs = 0
        self._n_vectors = 0
        self._header_extra = {
            'mode': '__MODE__',
            'sigma': '_sigma',
            'n_epochs': '_n_epochs',
            'n_vectors': '_n_vectors',
        }

    @property
    def sigma(self) -> float:
        """Starting value of the neighborhood radius."""
        return self._sigma

    # ------------------------------------------------------------------------------------------
    def train(self, dataset, dataset_weights=None, n_epochs: typing.Optional[int] = None,
              n_vectors: typing.Optional[int] = None, use_best_epoch: bool = True,
              stop_quantization_error: typing.Optional[float] = None,
              stop_topographic_error: typing.Optional[float] = None, show_progress_bar: bool = False,
              enable_gpu: bool = True, threads_per_blocks: typing.Optional[int] = None, *,
              precision: typing.Optional[str] = None, arithmetic: typing.Optional[str] = None) -> None:
        """Trains the neural network by number of epochs (``n_epochs``) or by number of vectors
        (``n_vectors``), exactly one of them (som_batch.py:214-505)."""
        if stop_quantization_error is None:
            stop_quantization_error = -1.0e6
        if stop_topographic_error is None:
            stop_topographic_error = -1.0e6
        self._n_epochs = n_epochs
        self._n_vectors = n_vectors
        if not ((n_epochs is not None and n_vectors is None) or (n_epochs is None and n_vectors is not None)):
            raise ValueError('Invalid training method, specify either `n_epochs` or `n_vectors`')
        self._require_gpu(enable_gpu)
        prec, arith = self._prec_codes(precision, arithmetic)
        sigma0 = self._device_dtype()(self._sigma)  # cast to the dtype, as the reference does (som_batch.py:315)
        mn = self._m * self._n
        model = self._device_model()
        ctx = self._ctx()
        acc = Accumulator(model)

        if n_epochs is not None:
            self._history = np.empty((n_epochs, mn, self._dim), dtype=self._dtype)
            self._quantization_errors = np.full(n_epochs, np.nan, dtype=np.float64)
            self._topographic_errors = np.full(n_epochs, np.nan, dtype=np.float64)
            if n_epochs > 0:
                self._train_epochs(model, acc, ctx, dataset, dataset_weights, n_epochs, sigma0, prec, arith,
                                   stop_quantization_error, stop_topographic_error, show_progress_bar)
        else:
            self._history = np.empty((1, mn, self._dim), dtype=self._dtype)
            self._quantization_errors = np.zeros(1, dtype=np.float64)
            self._topographic_errors = np.zeros(1, dtype=np.float64)
            self._train_vectors(model, acc, ctx, dataset, dataset_weights, n_vectors, sigma0, prec, arith,
                                show_progress_bar)
        acc.close()
        if use_best_epoch:
            self._weights[:] = self._history[np.nanargmin(self._quantization_errors)]

    # ------------------------------------------------------------------------------------------
    def _errors_from(self, e3):
        if self._m * self._n < 2:
            return np.nan, np.nan
        return (e3[0] / e3[2], e3[1] / e3[2]) if e3[2] > 0.0 else (np.nan, np.nan)

    def _train_epochs(self, model, acc, ctx, dataset, dataset_weights, n_epochs, sigma0, prec, arith, stop_q,
                      stop_t, show_progress_bar):
        mn, dim = self._m * self._n, self._dim
        early_stop = stop_q > -1.0e6 or stop_t > -1.0e6
        history = DeviceArray.empty(ctx, (n_epochs, mn, dim), self._device_dtype())
        errors = DeviceArray.zeros(ctx, (n_epochs, 3), np.float64)
        resident = not callable(dataset) and (dataset_weights is None or not callable(dataset_weights))
        res_x = res_r = None
        if resident:
            x = np.ascontiguousarray(dataset, dtype=self._dtype)
            if x.ndim != 2 or x.shape[1] != dim:
                raise ValueError(f'dataset must have shape (N, {dim})')
            if dataset_weights is not None and np.shape(dataset_weights)[0] != x.shape[0]:
                raise ValueError('`dataset` and `dataset_weights` chunks must be aligned')
            res_x = DeviceArray.from_host(ctx, x)
            res_r = None if dataset_weights is None else DeviceArray.from_host(
                ctx, np.asarray(dataset_weights, dtype=np.float64).reshape(-1))
            n_res = x.shape[0]

        def chunks():
            if resident:
                yield res_x.ptr, (None if res_r is None else res_r.ptr), n_res
            else:
                for v, w in iterate_chunks(dataset, dataset_weights):
                    xp, rp, n = self._upload_chunk(v, w)
                    yield xp, rp, n

        epochs = range(n_epochs)
        if show_progress_bar:
            import tqdm
            epochs = tqdm.tqdm(epochs)
        run = 0
        for e in epochs:
            acc.reset()
            fused = None if (early_stop or e == 0) else errors.view_ptr((e - 1) * 3)
            for xp, rp, n in chunks():
                acc.epoch_accumulate(xp, rp, n, prec, arith, fused)
            acc.epoch_update(e, n_epochs, sigma0, history.view_ptr(e * mn * dim))
            run = e + 1
            if early_stop:
                eptr = errors.view_ptr(e * 3)
                for xp, rp, n in chunks():
                    model.compute_errors(xp, rp, n, eptr, prec, arith)
                e3 = errors.copy_to_host()[e]
                q, t = self._errors_from(e3)
                self._quantization_errors[e], self._topographic_errors[e] = q, t
                if q <= stop_q and t <= stop_t:
                    from .logging import logger
                    logger.info('Stopping at epoch #{}.'.format(e))
                    break
        if not early_stop:
            eptr = errors.view_ptr((n_epochs - 1) * 3)
            for xp, rp, n in chunks():
                model.compute_errors(xp, rp, n, eptr, prec, arith)
            e3 = errors.copy_to_host()
            for e in range(n_epochs):
                self._quantization_errors[e], self._topographic_errors[e] = self._errors_from(e3[e])
        h = history.copy_to_host()
        self._history[:run] = h[:run]
        self._weights[:] = self._history[run - 1]
        for a in (history, errors, res_x, res_r):
            if a is not None:
                a.free()

    def _train_vectors(self, model, acc, ctx, dataset, dataset_weights, n_vectors, sigma0, prec, arith,
                       show_progress_bar):
        mn, dim = self._m * self._n, self._dim
        acc.reset()
        cur = 0
        bar = None
        if show_progress_bar:
            import tqdm
            bar = tqdm.tqdm(total=n_vectors)
        for v, w in iterate_chunks(dataset, dataset_weights):
            count = min(v.shape[0], n_vectors - cur)
            if count <= 0:
                break
            xp, rp, _ = self._upload_chunk(v[:count], None if w is None else w[:count])
            acc.iter_accumulate(xp, rp, count, cur, n_vectors, sigma0, prec, arith)
            cur += count
            if bar is not None:
                bar.update(count)
            if cur >= n_vectors:
                break
        history = DeviceArray.empty(ctx, (mn, dim), self._device_dtype())
        acc.iter_update(history.ptr)
        self._history[0] = history.copy_to_host()
        self._weights[:] = self._history[0]
        history.free()
        q, t = self.compute_errors(dataset, dataset_weights, enable_gpu=True,
                                   precision=self._precision_name(prec), arithmetic=self._arith_name(arith))
        self._quantization_errors[0] = q
        self._topographic_errors[0] = t

    @staticmethod
    def _precision_name(code):
        return {v: k for k, v in _lib.PRECISIONS.items()}[code]

    @staticmethod
    def _arith_name(code):
        return {v: k for k, v in _lib.ARITHMETICS.items()}[code]
