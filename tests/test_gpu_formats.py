"""Formats on the GPU: a case built from a reference CASE directory steps
exactly like the generated case, and a restart written by one context and
read by another continues bit for bit."""
import numpy as np
import pytest

from beta1_amd import CaseDirDesc, HostCase, MercuryDesc, MhdContext, WindDesc, case_parts

pytestmark = pytest.mark.gpu

FIELDS = ("U_H", "B_xi", "B_eta", "B_zeta")


def fields(ctx, cid, names=FIELDS):
    nb = ctx.info(cid)[0]
    return {(f, b): ctx.field(cid, f, b) for f in names for b in range(nb)}


def test_case_from_dir_steps_like_generated_case(tmp_path):
    desc = WindDesc(nang=8, nrad=16, v_kms=380.0)
    d = str(tmp_path / "CASE")
    HostCase(desc).write_dir(d)
    ctx = MhdContext()
    a = ctx.add(desc)
    b = ctx.add(CaseDirDesc(d, desc))
    ctx.finalize()
    ctx.step(6)
    fa, fb = fields(ctx, a), fields(ctx, b)
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k
    assert ctx.last_dt(a) == ctx.last_dt(b)


@pytest.mark.parametrize("desc,names", [(WindDesc(nang=8, nrad=16), FIELDS),
                                        (MercuryDesc(nang=8, nrad=16), FIELDS + ("U_Na",))])
def test_restart_continues_bitwise(tmp_path, desc, names):
    path = str(tmp_path / "rst.bin")
    one = MhdContext()
    c1 = one.add(desc)
    one.finalize()
    one.step(5)
    one.write_restart(c1, path, step=5, time=0.25)
    two = MhdContext()
    c2 = two.add(desc)
    two.finalize()
    two.step(2)  # a different state, overwritten by the restart
    assert two.read_restart(c2, path) == (5, 0.25)
    ref = fields(one, c1, names)
    got = fields(two, c2, names)
    for k in ref:
        assert np.array_equal(ref[k], got[k]), k
    one.step(4)
    two.step(4)
    ref, got = fields(one, c1, names), fields(two, c2, names)
    for k in ref:
        assert np.array_equal(ref[k], got[k]), k
    assert one.last_dt(c1) == two.last_dt(c2)


@pytest.mark.parametrize("desc", [WindDesc(nang=8, nrad=16), MercuryDesc(nang=8, nrad=16)])
def test_tecplot_of_the_device_state_equals_host_writer(tmp_path, desc):
    """The device context's .plt after steps is byte-identical to the host
    writer's on the same state (downloaded into a host case)."""
    ctx = MhdContext()
    cid = ctx.add(desc)
    ctx.finalize()
    ctx.step(4)
    dev = str(tmp_path / "dev.plt")
    ctx.write_tecplot(cid, dev, time=0.75)
    hc = HostCase(desc)
    names = ("U_H", "B_xi", "B_eta", "B_zeta", "B_cell", "J_cell") + (("U_Na",) if isinstance(desc, MercuryDesc) else ())
    for (f, b), a in fields(ctx, cid, names).items():
        hc.set_field(f, b, a)
    host = str(tmp_path / "host.plt")
    hc.write_tecplot(host, time=0.75)
    assert open(dev, "rb").read() == open(host, "rb").read()


@pytest.mark.parametrize("layout,source,curved", [
    ("aligned", "dir", False), ("turned", "dir", False), ("aligned", "parts", False), ("turned", "parts", False),
    ("turned", "dir", True)])
def test_faces_shared_by_several_patches_match_the_oracle(tmp_path, layout, source, curved):
    """A CASE directory whose block faces are split among neighbours (the
    periodic box as three blocks, A's two i faces each shared by B and C, C
    aligned or turned half a turn about k) steps on the device bit for bit
    as the oracle steps it: the op lists the host builds from the patches
    and the oracle's own halo resolution agree, through the one-pass EMF.
    The case comes either from the directory or as a driver's arrays
    (mhd_case_begin .. mhd_set_face_parts .. mhd_case_end); the curved grid
    is the skewed, non-orthogonal variant of the box."""
    import test_face_patches as T
    from beta1_amd.fields import interior
    from oracle_bind import Oracle

    T.use(layout)
    three = T.three_block_dir(tmp_path, T.one_block_dir(tmp_path, curved))
    orc = Oracle(HostCase(CaseDirDesc(three, WindDesc()), ghost=2))
    T.fill(orc, [0, 1, 2])
    ctx = MhdContext()
    if source == "dir":
        cid = ctx.add(CaseDirDesc(three, WindDesc()))
    else:
        cid = ctx.add_parts(*case_parts(HostCase(CaseDirDesc(three, WindDesc()), ghost=2)))
    ctx.finalize()
    for b in range(3):
        for name in ("U_H",) + T.FACES:
            ctx.set_field(cid, name, b, orc.view(name, b))
    ctx.refresh()
    dts = orc.step(4)
    ctx.step(4)
    assert ctx.last_dt(cid) == dts[-1]
    for b, n in enumerate(T.DIMS):
        for name in ("U_H",) + T.FACES:
            sl = interior(name, n, 2)
            assert np.array_equal(ctx.field(cid, name, b)[sl], orc.view(name, b)[sl]), (name, b)
