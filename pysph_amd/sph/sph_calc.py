"""SPHCalc: one operation on one destination array from its sources
(reference: sph/sph_calc.pyx:27-276).

`sph(out1, out2, out3)` writes the named destination properties with the
operator's sum over all sources, computed on the GPU in one launch (the
reference zeroes the outputs, then accumulates one source at a time).
"""
import logging

logger = logging.getLogger(__name__)


class SPHCalc:
    def __init__(self, particles, sources, dest, kernel, funcs, updates, integrates=False,
                 dnum=0, nbr_info=True, id="", kernel_gradient_correction=False,
                 kernel_correction=-1, dim=1, snum=""):
        if kernel_correction != -1:
            raise NotImplementedError("kernel correction is not supported")
        self.nbr_info = nbr_info
        self.particles = particles
        self.sources = list(sources)
        self.nsrcs = len(self.sources)
        self.dest = dest
        self.nbr_locators = []
        self.nnps_manager = particles.nnps_manager
        self.cell_manager = particles.cell_manager
        self.funcs = list(funcs)
        self.kernel = kernel
        self.integrates = integrates
        self.updates = list(updates)
        self.nupdates = len(self.updates)
        self.kernel_correction = kernel_correction
        self.dnum = dnum
        self.id = id
        self.dim = dim
        self.snum = snum
        self.correction_manager = None
        self.tag = ""
        self.src_reads = []
        self.dst_reads = []
        self.initial_props = []
        self.dst_writes = {}
        self.check_internals()
        self.setup_internals()

    def check_internals(self):
        """Consistency checks of sph_calc.pyx:132-183."""
        if (len(self.sources) == 0 or self.nnps_manager is None or self.dest is None
                or self.kernel is None or len(self.funcs) == 0):
            logger.debug("calc %s has no sources", self.id)
            return
        if len(self.funcs) != len(self.sources):
            raise ValueError("One sph function is needed per source")
        funcs = self.funcs
        for i in range(len(funcs) - 1):
            if type(funcs[i]) != type(funcs[i + 1]):
                raise ValueError("All sph_funcs should be of same type")
            if funcs[i].tag != funcs[i + 1].tag:
                raise ValueError("All functions should have the same tag")
        for i, f in enumerate(funcs):
            if f.source is not self.sources[i]:
                raise ValueError("SPHFunction.source not same as SPHCalc.sources[%d]" % i)
            if f.num_outputs != len(self.updates):
                raise ValueError("Number of updates not same as num_outputs; required %d, "
                                 "got %d %s for func %s" % (f.num_outputs, len(self.updates),
                                                            self.updates, f))

    def setup_internals(self):
        self.nbr_locators = []
        self.tag = self.funcs[0].tag
        for src, func in zip(self.sources, self.funcs):
            func.kernel = self.kernel
            loc = self.nnps_manager.get_neighbor_particle_locator(src, self.dest,
                                                                  self.kernel.radius())
            func.nbr_locator = loc
            self.nbr_locators.append(loc)
        func = self.funcs[0]
        self.src_reads = func.src_reads
        self.dst_reads = func.dst_reads

    def sph(self, output_array1=None, output_array2=None, output_array3=None,
            exclude_self=False):
        """Evaluate into the named destination properties.  `exclude_self` is
        accepted and ignored, as in sph_calc.pyx:244-268; the functions' own
        exclude_self drops the self pair."""
        names = [n for n in (output_array1, output_array2, output_array3) if n is not None]
        # bin first: a re-binning permutes (replaces) the property tensors
        if self.nbr_info:
            self.cell_manager.update()
        outs = []
        for n in names:
            t = self.dest.get_carray(n)
            if t is None:
                raise AttributeError("property %s not present on %s" % (n, self.dest.name))
            outs.append(t)
        type(self.funcs[0]).calc_eval(self, outs)
        if self.dest.is_dirty:
            self.particles.update()

    def sph_array(self, output1, output2=None, output3=None, exclude_self=False):
        """Evaluate into device tensors instead of named properties
        (`exclude_self` ignored, as in sph())."""
        outs = [o for o in (output1, output2, output3) if o is not None]
        if self.nbr_info:
            self.cell_manager.update()
        type(self.funcs[0]).calc_eval(self, outs)
