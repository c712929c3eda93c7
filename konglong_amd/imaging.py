"""Imaging conditions on the GPU: the zero-lag correlation with its tapers and illumination
(ImagingCorrelator::crossone + ImagingCondition::calIllum) and ImagingCorrelator's FFT options
(LowcutFilter Laplacian/none, Inversion, Inversion2, and all three with the raw image)."""
import ctypes as C

from ._lib import lib, check, IC_IMAGE1, IC_INVERSION, IC_INVERSION2, IC_ALLIN3  # noqa: F401


def _p(d):
    return None if d is None else d.ptr


def correlate(src, rec, nx, ny, nz, taper_t=1.0, taper_z=None, ntaper_z=0, image=None, illum=None, image64=None,
              illum64=None):
    """image += S*R, illum += S*S with S = src*taper (device arrays; outputs optional)."""
    check(lib().klrtm_ic_correlate(src.ptr, rec.ptr, nx, ny, nz, taper_t, _p(taper_z), ntaper_z, _p(image), _p(illum),
                                   _p(image64), _p(illum64)))


class ImagingCorrelator:
    """ImagingCorrelator(nx, ny, nz, dx, dy, dz, nx2=nx, ny2=ny, nz2=nz, kmin) (bgprtm/ImagingCorrelator.h:20-33)
    (set_double for nDbl = 2); device volumes (ny, nx, nz).  run() takes the option of ImagingCondition::getImagingOption;
    src and rec are work buffers (INVERSION2 filters them in place, as the reference does)."""

    def __init__(self, nx, ny, nz, dx, dy, dz, kmin=1.0 / 1492.0):
        self.h = C.c_void_p()
        self.shape = (ny, nx, nz)
        check(lib().klrtm_correlator_create(C.byref(self.h), nx, ny, nz, dx, dy, dz, kmin))

    def run(self, src, rec, option, image=None, img_fwi=None, img_raw=None, image64=None):
        check(lib().klrtm_correlator_run(self.h, _p(image), _p(img_fwi), _p(img_raw), _p(image64), src.ptr, rec.ptr,
                                         option))

    def costerm(self, image):
        check(lib().klrtm_correlator_costerm(self.h, image.ptr))

    def sinterm(self, imgfull, imgcos):
        check(lib().klrtm_correlator_sinterm(self.h, _p(imgfull), imgcos.ptr))

    def unit_derivative(self, src, axis, out):
        check(lib().klrtm_correlator_unit_derivative(self.h, src.ptr, axis, out.ptr))

    def kfilter(self, wave):
        check(lib().klrtm_correlator_kfilter(self.h, wave.ptr))

    def set_double(self, dz=2, dx=2, dy=2, high_z=0.9, high_xy=0.7):
        """The doubled imaging grid (nx2 = dx nx, ny2 = dy ny, nz2 = dz nz of the reference's constructor):
        run() then fills nloop sub-images per image; interleave() puts them onto the doubled grid."""
        check(lib().klrtm_correlator_set_double(self.h, int(dz), int(dx), int(dy), float(high_z), float(high_xy)))
        self.dbl = (int(dz), int(dx), int(dy))

    @property
    def nloop(self):
        n = C.c_int()
        check(lib().klrtm_correlator_nloop(self.h, C.byref(n)))
        return n.value

    def interleave(self, sub, out):
        check(lib().klrtm_correlator_interleave(self.h, sub.ptr, out.ptr))

    def run_copies(self, src, rec, option, image=None, img_fwi=None):
        """ImagingCorrelator::run on nloop-volume buffers (applyFFTV, and for IMAGE1 interp, done by the
        caller)."""
        check(lib().klrtm_correlator_run_copies(self.h, _p(image), _p(img_fwi), src.ptr, rec.ptr, int(option)))

    def interp(self, buf):
        """ImagingCorrelator::interp: copies 1.. of buf (nloop volumes) from copy 0."""
        check(lib().klrtm_correlator_interp(self.h, buf.ptr))

    def hightape(self, wave):
        """ImagingCondition::applyFFTV in place."""
        check(lib().klrtm_correlator_hightape(self.h, wave.ptr))

    def interp_illum(self, ill, out):
        """ImagingCondition::interpIllum onto the doubled grid."""
        check(lib().klrtm_correlator_interp_illum(self.h, ill.ptr, out.ptr))

    def close(self):
        if self.h:
            lib().klrtm_correlator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
