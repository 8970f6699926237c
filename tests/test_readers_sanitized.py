"""The host readers (.control files, SpecMesh meshes, .hsol files) built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++, host code only) and driven over the fixtures and over malformed
inputs: empty and truncated files, keys and values longer than any buffer, unterminated quotes,
incomplete #define blocks, numbers that do not parse, missing files.  Each good input must read,
each malformed one must fail with an error code, and the sanitizers must report nothing."""
import ast
import os
import shutil
import subprocess

import numpy as np
import pytest

from horses3d_amd import mesh, specmesh

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CTL = os.path.join(HERE, "golden", "control")
HSOL = os.path.join(HERE, "golden", "hsol")
CASES = ("order_file_euler_pirozzoli", "orders_ijk_smagorinsky", "syntax_split_kg_lxf")
SOLS = ("box2_mixed_p2to4", "box2_p3_curved")
BAD_CONTROL = {
    "empty": "",
    "noeq": "just some words\nand more\n",
    "longkey": ("k" * 300) + " = 1\n" + "mesh file name = " + "m" * 5000 + "\n",
    "quote": 'mesh file name = "unterminated\nmach number = "0.1\n',
    "define": "#define boundary\n#define boundary a__b__c\n type = periodic\n coupled boundary\n#end\n"
              "#define volume monitor\n",
    "eqonly": "=\n = \n==\n!\n/\n\n   \n\t=\t\n",
    "nomesh": "mach number = 0.1\nreynolds number = 100\npolynomial order = 3\n",
    "badnum": "mesh file name = box.mesh\nmach number = abc\nreynolds number = 1e\npolynomial order = x\ncfl = \n",
}
MESH_CUTS = ("0", "0.01", "0.1", "0.5", "0.9", "0.999")
HSOL_CUTS = ("0", "0.05", "0.5", "0.99")


def _stage(d):
    for name in CASES:
        c = os.path.join(d, "case_" + name)
        os.makedirs(c)
        z = np.load(os.path.join(CTL, name + ".npz"), allow_pickle=False)
        meta = ast.literal_eval(str(z["meta"]))
        shutil.copy(os.path.join(CTL, name + ".control"), os.path.join(c, "case.control"))
        if os.path.exists(os.path.join(CTL, name + "_orders.txt")):
            shutil.copy(os.path.join(CTL, name + "_orders.txt"), os.path.join(c, "orders.txt"))
        specmesh.write_specmesh(os.path.join(c, "box.mesh"), meta["n"], meta["amp"],
                                mesh.box_phases(meta["phase_seed"]), bface_order=meta["bface_order"])
    bad = os.path.join(d, "bad")
    os.makedirs(bad)
    for k, v in BAD_CONTROL.items():
        with open(os.path.join(bad, k + ".control"), "w") as f:
            f.write(v)
    data = open(os.path.join(d, "case_" + CASES[0], "box.mesh"), "rb").read()
    for cut in MESH_CUTS:
        with open(os.path.join(bad, "trunc_%s.mesh" % cut), "wb") as f:
            f.write(data[:int(len(data) * float(cut))])
    for s in SOLS:
        z = np.load(os.path.join(HSOL, s + ".npz"), allow_pickle=False)
        with open(os.path.join(d, s + ".N"), "w") as f:
            f.write(" ".join(str(int(x)) for x in z["N"].ravel()))
        for part in ("_step0", "_final"):
            b = open(os.path.join(HSOL, s + part + ".hsol"), "rb").read()
            for cut in HSOL_CUTS:
                with open(os.path.join(bad, "trunc_%s_%s%s.hsol" % (cut, s, part)), "wb") as f:
                    f.write(b[:int(len(b) * float(cut))])


def _build(d):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    flags = ["-std=c++17", "-g", "-O1", "-fsanitize=address,undefined", "-fno-omit-frame-pointer"]
    src = [os.path.join(HERE, "sanitize", "readers.cpp")] + [
        os.path.join(REPO, "horses3d_amd", "csrc", f) for f in ("control.cpp", "specmesh.cpp", "hsol.cpp")]
    objs = [os.path.join(d, "o%d.o" % i) for i in range(len(src))]
    procs = [subprocess.Popen([gxx] + flags + ["-c", s, "-o", o], stderr=subprocess.PIPE, text=True)
             for s, o in zip(src, objs)]
    for p in procs:
        err = p.communicate()[1]
        assert p.returncode == 0, err
    exe = os.path.join(d, "readers")
    r = subprocess.run([gxx] + flags + objs + ["-o", exe], capture_output=True, text=True)
    if r.returncode != 0 and "asan" in r.stderr.lower():
        pytest.skip("sanitizer runtime not available: " + r.stderr[-200:])
    assert r.returncode == 0, r.stderr
    return exe


def test_readers_under_sanitizers(tmp_path):
    d = str(tmp_path)
    _stage(d)
    exe = _build(d)
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([exe, d, HSOL], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    out = r.stdout.splitlines()
    assert out[-1] == "done"

    def lines(prefix):
        return [ln for ln in out if ln.startswith(prefix)]

    # good control files read, and their cases build from the case folder
    for name in CASES:
        ctl = lines("control %s/case_%s/case.control" % (d, name))
        assert len(ctl) == 2 and all(" rc 0" in ln for ln in ctl)
    cases = lines("  case rc")
    assert [ln.split()[2] for ln in cases[:6:2]] == ["0", "0", "0"]          # with the case folder
    assert all(ln.split()[2] == "1" for ln in cases[1:6:2])                 # relative names, wrong folder
    assert all(ln.split()[2] != "0" for ln in cases[6:])                    # every malformed file
    # good meshes read with their periodic pairs; truncated ones fail
    for name in CASES:
        assert " rc 0" in lines("mesh %s/case_%s/box.mesh [left=right;front=back;bottom=top]" % (d, name))[0]
    for cut in MESH_CUTS:
        got = lines("mesh %s/bad/trunc_%s.mesh" % (d, cut))
        assert got and all(" rc 0" not in ln for ln in got), got
    # solution files: fixtures read and round-trip; truncated bodies and headers fail
    for s in SOLS:
        for part in ("_step0", "_final"):
            i = out.index(next(ln for ln in out if ln.startswith("hsol %s/%s%s.hsol" % (HSOL, s, part))))
            assert " rc 0 " in out[i] and out[i + 1] == "  read rc 0 " and out[i + 3].endswith("equal 1")
            for cut in HSOL_CUTS:
                j = out.index(next(ln for ln in out if ln.startswith("hsol %s/bad/trunc_%s_%s%s" % (d, cut, s, part))))
                assert " rc 1 " in out[j] or out[j + 1].startswith("  read rc 1"), out[j:j + 2]
