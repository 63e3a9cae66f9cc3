"""NISTGPUDropIn: the device Newton step behind the reference's own solver plugin contract, on
the reference's own objects.

The reference calls a solver as ``Solver(jobInfo, journal, **options)`` and
``solveStep(step, model, fieldOutputController, outputmanagers)``
(edelweissfe/solvers/nonlinearimplicitstatic.py:82-110, looked up by name in
config/solvers.py:48-76).  This class takes exactly those arguments:

* ``model``: the reference's FEModel (models/femodel.py:45-298): ``nodeFields["displacement"]``
  (a NodeField, fields/nodefield.py:83-90, whose ``nodes`` order is the DOF order of
  DofManager, numerics/dofmanager.py:259-300), ``elements`` (dict, element order = VIJ order),
  ``nodeSets``, ``elementSets``, ``surfaces`` and ``time``.  Element types are read from the
  DisplacementElement (``nNodes`` and its quadrature point count, elements/library.py:212-291),
  materials per element from the element's material object (``VonMisesMaterial``,
  ``LinearElasticMaterial``, or any object with ``materialProperties`` whose class name names
  the material), so models with several sections (sections/base/sectionbase.py:86-128) and
  per-element properties (materialParameterFromField) come through as they are.
* ``step``: the reference's AdaptiveStep (steps/adaptivestep.py:65-160), used as it is
  (``getTimeStep``, ``discardAndChangeIncrement``, ``preventIncrementIncrease``, ``maxIter``,
  ``criticalIter``, ``maxGrowIter``, ``cutbackFactor``, ``number``), with its step actions:
  dirichlet (stepactions/dirichlet.py:53-112: ``nSet`` of nodes, ``components``,
  ``getDelta(timeStep)``), nodeforces (nodeforces.py:53-152: ``nodeSet``,
  ``getCurrentLoad(timeStep)`` per node), bodyforce (bodyforce.py: ``elementSet``,
  ``getCurrentLoad``), distributedload (distributedload.py:52-110: ``surface``, ``loadType``,
  ``getCurrentLoad``), geostatic (geostatic.py:51-101: ``geostaticElements``,
  ``geostaticDefinition``, ``active``) and ``options["NISTSolver"].options``
  (stepactions/options.py).
* ``fieldOutputController`` and ``outputmanagers`` as the reference passes them.

The device model (mesh, pattern, QP state on the GPU) is kept across the steps of one
reference model, so the state of one step (a geostatic prestress, plastic strains) is the
start of the next, as the reference's elements keep theirs.  After every converged increment
the node fields U, P and dU are written back when a field output controller or output managers
read them; at the end of the step always, and the reference elements receive the accepted QP
state (``_stateVarsRef``, element.py:222-240) and the model its time.
"""

import numpy as np

from ..model import FEModel as DeviceModel
from .nistgpu import NISTGPU

# reference solver names that select a direct factorisation; the device replaces them by its
# Krylov solvers (PCG, or GMRES for a non-symmetric material)
_DIRECT = ("pardiso", "superlu", "klu", "umfpack", "mumps", "spsolve", "")


def element_type_of(el):
    """C3D8 / C3D20 / C3D20R from a reference element (DisplacementElement stores nNodes and
    its quadrature table, elements/library.py:212-291)."""
    name = getattr(el, "elType", None) or getattr(el, "elementType", None)
    if isinstance(name, str):
        key = name.upper().rstrip("N") if name.upper() in ("C3D8N", "C3D20N") else name.upper()
        if key in ("C3D8", "C3D20", "C3D20R"):
            return key
    dims = getattr(el, "nSpatialDimensions", 3)
    nn = el.nNodes
    nint = getattr(el, "_nInt", None)
    if dims == 3 and nn == 8 and nint in (8, None):
        return "C3D8"
    if dims == 3 and nn == 20 and nint in (27, None):
        return "C3D20"
    if dims == 3 and nn == 20 and nint == 8:
        return "C3D20R"
    raise NotImplementedError("element {} ({} nodes, {} QPs) has no device counterpart".format(
        type(el).__name__, nn, nint))


def material_of(mat):
    """(name, properties) of a reference material object: VonMisesMaterial keeps its parameters
    as attributes (vonmises.py:105-113), LinearElasticMaterial too (linearelastic.py:59-61);
    other materials expose materialProperties."""
    cls = type(mat).__name__.lower()
    if "vonmises" in cls:
        if hasattr(mat, "materialProperties"):
            return "vonmises", np.asarray(mat.materialProperties, dtype=float)[:6]
        return "vonmises", np.array([mat._E, mat._v, mat.yieldStress, mat.HLin, mat.deltaYieldStress, mat.delta])
    if "linearelastic" in cls:
        if hasattr(mat, "materialProperties"):
            return "linearelastic", np.asarray(mat.materialProperties, dtype=float)[:2]
        return "linearelastic", np.array([mat._E, mat._v])
    if "druckerprager" in cls and hasattr(mat, "materialProperties"):
        return "druckerprager", np.asarray(mat.materialProperties, dtype=float)[:6]
    raise NotImplementedError("material {} has no device counterpart".format(type(mat).__name__))


class ModelMap:
    """The device model of a reference model and the maps between them."""

    def __init__(self, model):
        field = model.nodeFields["displacement"]
        self.field = field
        self.nodes = list(field.nodes)
        self.nodePos = {n: i for i, n in enumerate(self.nodes)}
        self.elements = list(model.elements.values())
        self.elementPos = {el: i for i, el in enumerate(self.elements)}
        coords = np.array([np.asarray(n.coordinates, dtype=float)[:3] for n in self.nodes])
        types = {element_type_of(el) for el in self.elements}
        if len(types) != 1:
            raise NotImplementedError("elements of several types in one model: {}".format(sorted(types)))
        self.elType = types.pop()
        conn = np.array([[self.nodePos[n] for n in el.nodes] for el in self.elements], dtype=np.int64)
        mats = [material_of(el.material) for el in self.elements]
        self.dm = DeviceModel(coords, self.elType, conn, {}, {})
        names = [m[0] for m in mats]
        if len(set(names)) == 1 and all(np.array_equal(m[1], mats[0][1]) for m in mats):
            self.dm.assignMaterial(names[0], mats[0][1])
        else:  # sections, or per-element properties: one row of 6 per element
            props = np.zeros((len(mats), 6))
            for e, (_, p) in enumerate(mats):
                props[e, :p.size] = p
            if len(set(names)) == 1:
                self.dm.assignMaterial(names[0], props[:, :mats[0][1].size])
            else:
                self.dm.assignMaterial(np.array(names), props)
        self._sets = {}
        self.fingerprint = self._fingerprint(model)

    def _fingerprint(self, model):
        return (id(model), len(self.nodes), len(self.elements))

    def nodeSet(self, nodes):
        """Registers the node positions of a reference node set; returns its device name."""
        key = ("n", id(nodes))
        if key not in self._sets:
            name = "_ref_nset_%d" % len(self._sets)
            self.dm.nodeSets[name] = np.array([self.nodePos[n] for n in nodes], dtype=np.int64)
            self._sets[key] = (name, nodes)  # the set kept alive: its id stays its own
        return self._sets[key][0]

    def elementSet(self, elements):
        key = ("e", id(elements))
        if key not in self._sets:
            name = "_ref_elset_%d" % len(self._sets)
            self.dm.elementSets[name] = np.array([self.elementPos[el] for el in elements], dtype=np.int64)
            self._sets[key] = (name, elements)  # the set kept alive: its id stays its own
        return self._sets[key][0]

    def surface(self, surface):
        key = ("s", id(surface))
        if key not in self._sets:
            name = "_ref_surface_%d" % len(self._sets)
            self.dm.surfaces[name] = {int(f): np.array([self.elementPos[el] for el in els], dtype=np.int64)
                                      for f, els in surface.items()}
            self._sets[key] = (name, surface)  # the set kept alive: its id stays its own
        return self._sets[key][0]

    def readField(self, name):
        return np.asarray(self.field[name], dtype=float).reshape(-1)

    def writeFields(self):
        for name, v in (("U", self.dm.U), ("P", self.dm.P), ("dU", self.dm.dU)):
            try:
                dst = self.field[name]
            except (KeyError, AttributeError):
                continue
            dst[...] = np.asarray(v).reshape(np.shape(dst))


class _Dirichlet:
    """dirichlet.StepAction seen by NISTGPU: DOFs 3 p + c of its node set and components."""

    def __init__(self, ref, mm, refModel):
        self.ref, self.refModel = ref, refModel
        pos = np.array([mm.nodePos[n] for n in ref.nSet], dtype=np.int64)
        comps = np.asarray(ref.components, dtype=np.int64)
        self._idx = (3 * pos[:, None] + comps[None, :]).ravel()

    def indices(self, model):
        return self._idx

    def getDelta(self, model, timeStep):
        return np.asarray(self.ref.getDelta(timeStep), dtype=float).ravel()

    def applyAtStepEnd(self, model):
        self.ref.applyAtStepEnd(self.refModel)


class _Load:
    """nodeforces / bodyforce / distributedload StepAction seen by NISTGPU."""

    def __init__(self, ref, refModel, **names):
        self.ref, self.refModel = ref, refModel
        for k, v in names.items():
            setattr(self, k, v)
        self.loadType = getattr(ref, "loadType", None)

    def getCurrentLoad(self, timeStep):
        return np.asarray(self.ref.getCurrentLoad(timeStep), dtype=float)

    def applyAtStepEnd(self, model):
        self.ref.applyAtStepEnd(self.refModel)


class _Geostatic:
    """geostatic.StepAction: its definition set into the device state of its elements at every
    iteration start while active; at step end the reference's own action resets U (and the
    device model's U follows)."""

    def __init__(self, ref, mm, refModel):
        self.ref, self.refModel = ref, refModel
        els = getattr(ref, "geostaticElements", None)
        self.elements = None if els is None else np.array([mm.elementPos[el] for el in els], dtype=np.int64)
        if self.elements is not None and self.elements.size == len(mm.elements):
            self.elements = None

    def applyAtIterationStart(self, system, model):
        if not self.ref.active:
            return
        system.set_geostatic(np.asarray(self.ref.geostaticDefinition, dtype=float), self.elements)

    def applyAtStepEnd(self, model):
        if self.ref.active:
            model.U[:] = 0.0
        self.ref.applyAtStepEnd(self.refModel)


class _Step:
    """The reference step with its actions replaced by device views (the increment control is
    the reference step's own)."""

    def __init__(self, step, actions):
        self._step = step
        self.actions = actions

    def __getattr__(self, name):
        return getattr(self._step, name)


class _FieldOutput:
    """Writes the device results into the reference node fields before the reference field
    output controller records an increment."""

    def __init__(self, controller, mm):
        self.controller, self.mm = controller, mm

    def finalizeIncrement(self):
        self.mm.writeFields()
        if self.controller is not None:
            self.controller.finalizeIncrement()


class NISTGPUDropIn:
    identification = "NISTGPUSolver"

    def __init__(self, jobInfo=None, journal=None, **kwargs):
        self.device = NISTGPU(jobInfo, journal, **kwargs)
        self.journal = journal
        self._mm = None

    def _modelMap(self, model):
        if self._mm is not None and self._mm.fingerprint == (id(model), len(model.nodeFields["displacement"].nodes),
                                                             len(model.elements)):
            return self._mm
        self._mm = ModelMap(model)
        return self._mm

    def _adaptActions(self, step, model, mm):
        a = step.actions
        out = {"dirichlet": {}, "nodeforces": {}, "bodyforce": {}, "distributedload": {}, "geostatic": {},
               "options": {}}
        for k, d in a.get("dirichlet", {}).items():
            out["dirichlet"][k] = _Dirichlet(d, mm, model)
        for k, nf in a.get("nodeforces", {}).items():
            out["nodeforces"][k] = _Load(nf, model, nSet=mm.nodeSet(nf.nodeSet))
        for k, bf in a.get("bodyforce", {}).items():
            out["bodyforce"][k] = _Load(bf, model, elSet=mm.elementSet(bf.elementSet))
        for k, dl in a.get("distributedload", {}).items():
            t = str(dl.loadType).replace(" ", "").lower()
            if t != "surfacetraction":
                raise NotImplementedError("distributed load type {} on the device".format(dl.loadType))
            out["distributedload"][k] = _Load(dl, model, surface=mm.surface(dl.surface))
        for k, g in a.get("geostatic", {}).items():
            out["geostatic"][k] = _Geostatic(g, mm, model)
        opts = a.get("options", {}).get("NISTSolver", None)
        if opts is not None:
            opts = dict(getattr(opts, "options", opts))
            ls = str(opts.get("linsolver", "pcg")).lower()
            if ls in _DIRECT:  # the reference's direct solvers: the device Krylov default
                names = mm.dm.material["name"]
                dp = np.any(np.asarray(names) == "druckerprager")
                opts["linsolver"] = "gmres" if dp else "pcg"
            out["options"]["NISTSolver"] = opts
        unknown = set(a) - set(out)
        for k in unknown:
            if a[k]:
                raise NotImplementedError("step action group {} on the device".format(k))
        return out

    def solveStep(self, step, model, fieldOutputController=None, outputmanagers=()):
        mm = self._modelMap(model)
        dm = mm.dm
        dm.U = mm.readField("U").copy()
        dm.time = model.time
        dstep = _Step(step, self._adaptActions(step, model, mm))
        fo = _FieldOutput(fieldOutputController, mm)
        ok, _ = self.device.solveStep(dstep, dm, fo, outputmanagers)
        mm.writeFields()
        # the accepted QP state into the reference elements (element.py:222-240) and the time
        state = dm.qpState
        for el, st in zip(mm.elements, state):
            ref = getattr(el, "_stateVarsRef", None)
            # [QP][stress 6, strain 6, material state]: an element whose material has fewer
            # state variables than the widest of the model takes its leading columns
            if ref is not None and np.ndim(ref) == 2 and ref.shape[0] == st.shape[0] and ref.shape[1] <= st.shape[1]:
                ref[...] = st[:, :ref.shape[1]]
        model.time = dm.time
        return ok, model
