"""ImagingCorrelator::crossone + ImagingCondition::calIllum on the GPU (zero lag)."""
from ._lib import lib, check


def _p(d):
    return None if d is None else d.ptr


def correlate(src, rec, nx, ny, nz, taper_t=1.0, taper_z=None, ntaper_z=0, image=None, illum=None, image64=None,
              illum64=None):
    """image += S*R, illum += S*S with S = src*taper (device arrays; outputs optional)."""
    check(lib().klrtm_ic_correlate(src.ptr, rec.ptr, nx, ny, nz, taper_t, _p(taper_z), ntaper_z, _p(image), _p(illum),
                                   _p(image64), _p(illum64)))
