"""k-space kernels of the reference (fastpm/force/kernels.py:3-35) as device
operators.

Each operator is a product of factors evaluated per mode on the GPU:
``laplace`` (1/k^2, zero mode 0), ``gradient(dir, order)`` (i k_d, Nyquist
masked, or the 4-point i (8 sin w - sin 2w)/(6h)), ``longrange(r_split)``
(exp(-k^2 r^2)) and ``compensation(p)`` (divide by the order-p window,
sinc^p per axis).  Operators compose with ``*``: ``laplace * gradient(0, 1)``
is one fused multiplier, so a chain of ``apply`` calls in the reference becomes
one pass here.
"""
from .. import _lib


class KernelOp:
    def __init__(self, factor=1.0, laplace=False, grads=(), grad_order=1, r_split=0.0, deconv=0):
        if len(grads) > 2:
            raise ValueError("at most two gradient factors per operator")
        self.factor = float(factor)
        self.laplace = bool(laplace)
        self.grads = tuple(int(g) for g in grads)
        self.grad_order = int(grad_order)
        self.r_split = float(r_split)
        self.deconv = int(deconv)

    def __mul__(self, other):
        if isinstance(other, (int, float)):
            return KernelOp(self.factor * other, self.laplace, self.grads, self.grad_order,
                            self.r_split, self.deconv)
        if not isinstance(other, KernelOp):
            return NotImplemented
        if self.laplace and other.laplace:
            raise ValueError("a squared Laplacian is not supported")
        if self.grads and other.grads and self.grad_order != other.grad_order:
            raise ValueError("gradients of different orders cannot be fused")
        if self.r_split and other.r_split:
            raise ValueError("only one long-range split per operator")
        if self.deconv and other.deconv:
            raise ValueError("only one window compensation per operator")
        return KernelOp(self.factor * other.factor, self.laplace or other.laplace,
                        self.grads + other.grads,
                        self.grad_order if self.grads else other.grad_order,
                        self.r_split or other.r_split, self.deconv or other.deconv)

    __rmul__ = __mul__

    def fpm_kop(self):
        k = _lib.FpmKop()
        k.factor = self.factor
        k.r_split = self.r_split
        k.laplace = 1 if self.laplace else 0
        k.ngrad = len(self.grads)
        for i, g in enumerate(self.grads):
            k.grad_dir[i] = g
        k.grad_order = self.grad_order
        k.deconv = self.deconv
        return k

    def __repr__(self):
        return ("KernelOp(factor=%g, laplace=%s, grads=%s, order=%d, r_split=%g, deconv=%d)"
                % (self.factor, self.laplace, self.grads, self.grad_order, self.r_split,
                   self.deconv))


identity = KernelOp()
laplace = KernelOp(laplace=True)


def gradient(dir, order=0):
    """i a_dir(k): order 0 exact with the Nyquist plane cleared, order 1 the
    4-point finite difference (kernels.py:11-25)."""
    if order not in (0, 1):
        raise ValueError("gradient order must be 0 or 1")
    return KernelOp(grads=(dir,), grad_order=order)


def longrange(r_split):
    """Gaussian long-range split exp(-k^2 r^2); identity for r_split = 0
    (kernels.py:27-35)."""
    return KernelOp(r_split=r_split)


def compensation(order):
    """Divide by the mass-assignment window sinc(k h / 2)^order per axis
    (order 2: CIC, 3: TSC)."""
    return KernelOp(deconv=order)
