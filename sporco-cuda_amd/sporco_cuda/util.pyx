# cython: embedsignature=True, language_level=3
"""HIP device queries, drop-in for the reference ``sporco_cuda.util``
(sporco_cuda/util.pyx:1-45, backed by sporco_cuda/src/utils.cu:146-175)."""

cdef extern from "sporco_amd.h":
    int scc_get_device_count()
    int scc_get_current_device(int *dev)
    int scc_set_current_device(int dev)
    int scc_get_memory_info(size_t *free_bytes, size_t *total_bytes)
    const char *scc_get_device_name(int dev)


def device_count():
    """Number of HIP devices (0 when none, util.pyx:10-12)."""
    return scc_get_device_count()


def current_device(newdev=None):
    """Current device, after switching to ``newdev`` if given
    (util.pyx:15-26)."""
    cdef int dev
    if newdev is not None:
        dev = newdev
        err = scc_set_current_device(dev)
        if err != 0:
            raise RuntimeError('HIP error %d while setting device' % err)
    err = scc_get_current_device(&dev)
    if err != 0:
        raise RuntimeError('HIP error %d while getting current device' % err)
    return dev


def memory_info():
    """(free, total) device memory in bytes (util.pyx:29-35)."""
    cdef size_t free, total
    err = scc_get_memory_info(&free, &total)
    if err != 0:
        raise RuntimeError('HIP error %d' % err)
    return free, total


def device_name(int dev=0):
    """Device name, or None (util.pyx:38-45)."""
    cdef const char* cstr = scc_get_device_name(dev)
    if cstr == NULL:
        return None
    return (<bytes> cstr).decode('UTF-8')
