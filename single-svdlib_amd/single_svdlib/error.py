"""SvdLibError and its variants (src/error.rs:3-22).

Status codes of the C ABI map one-to-one onto the variants; shape and
argument errors, which the reference raises as panics (assert_eq!), are
ValueError here, allocation failures MemoryError, HIP failures RuntimeError.
"""


class SvdLibError(Exception):
    """Base of the reference's SvdLibError variants."""


class ImtqlbError(SvdLibError):
    pass


class StartvError(SvdLibError):
    pass


class StponeError(SvdLibError):
    pass


class Imtql2Error(SvdLibError):
    pass


class Las2Error(SvdLibError):
    pass


class NDArrayError(SvdLibError):
    pass


_BY_STATUS = {1: ImtqlbError, 2: StartvError, 3: StponeError, 4: Imtql2Error, 5: Las2Error, 6: NDArrayError}
_PREFIX = {1: "svdlibrs/imtqlb: ", 2: "svdlibrs/startv: ", 3: "svdlibrs/stpone: ", 4: "svdlibrs/imtql2: ",
           5: "svdlibrs/svdLas2: ", 6: "svdlibrs/ndarray: "}


def from_status(status, msg):
    if status in _BY_STATUS:
        return _BY_STATUS[status](_PREFIX[status] + msg)
    if status == 7:
        return ValueError(msg)
    if status == 8:
        return MemoryError(msg)
    if status == 10:
        return ValueError(msg)
    if status == 11:
        return RuntimeError("no HIP device: " + msg)
    return RuntimeError(f"svdh status {status}: {msg}")
