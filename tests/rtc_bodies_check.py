"""Compiles every user body the GPU tests use (module-level HIP C++ strings
defining struct UserLaw / UserNoise / UserRule / UserMedium) through
cbm_compile_check, for a model that instantiates its largest skeleton.
Run as a script (by test_rtc_bodies.py, in a child process); with --torch
it imports torch first, as the GPU test processes do, so the library binds
the HIP runtime and hipRTC that torch loaded.  Prints each body's name
before compiling it and exits 0 when all compile."""
import ctypes as C
import importlib
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT, os.path.join(ROOT, "cellbasedmodels.jl_amd")):
    sys.path.insert(0, p)

MODULES = ["law_bodies", "medium_bodies", "model_bodies", "rule_bodies", "test_api_host", "test_gpu_custom_law", "test_gpu_law_self",
           "test_gpu_sde_body", "test_gpu_user_de", "test_gpu_user_de_rules", "test_gpu_user_rule"]


def bodies():
    for name in MODULES:
        mod = importlib.import_module(name)
        for attr, v in sorted(vars(mod).items()):
            if isinstance(v, str) and "struct User" in v:
                yield f"{name}.{attr}", v


def check(A, label, body):
    from helpers import model
    log = C.create_string_buffer(1 << 16)
    args = [None, None, None, None]
    if "struct UserModel" in body:
        print(label, flush=True)
        rc = A.lib().cbm_compile_model_scope_check(body.encode(), log, len(log))
        if rc != A.OK:
            print(f"  FAILED ({rc}): {log.value.decode(errors='replace')[:2000]}", flush=True)
        return rc == A.OK
    if "struct UserMedium" in body:
        m = model(2, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, cell_edge=2.0)
        m.medium_n[0], m.medium_n[1] = 8, 8
        args[3] = body
    elif "struct UserRule" in body:
        m = model(3, neighbors=A.FULL, law=A.LAW_LINEAR, rule=A.RULE_CUSTOM)
        args[2] = body
    elif "struct UserNoise" in body:
        m = model(3, integrator=A.EM, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, cell_edge=2.0)
        args[1] = body
    else:
        m = model(3, neighbors=A.CELL_LINKED, law=A.LAW_CUSTOM, cell_edge=2.0)
        m.user_de = 4 if " ode(" in body else (1 if " de(" in body else 0)
        args[0] = body
    enc = [a.encode() if a is not None else None for a in args]
    print(label, flush=True)
    rc = A.lib().cbm_compile_check(C.byref(m), *enc, log, len(log))
    if rc != A.OK:
        print(f"  FAILED ({rc}): {log.value.decode(errors='replace')[:2000]}", flush=True)
    return rc == A.OK


def main():
    if "--torch" in sys.argv:
        import torch  # noqa: F401  the HIP runtime and hipRTC torch bundles
    import cbm._abi as A
    ok = [check(A, label, body) for label, body in bodies()]
    print(f"{sum(ok)} of {len(ok)} compiled", flush=True)
    return 0 if all(ok) and ok else 1


if __name__ == "__main__":
    sys.exit(main())
