"""Build the N-API addon swe_napi.node beside this file (needs Node's headers).

The addon links libswe_amd.so from the package directory through an $ORIGIN
rpath; the N-API symbols resolve against the running node binary at load time.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
NODE_INC = os.environ.get("NODE_INCLUDE", "/usr/include/node")


def build():
    out = os.path.join(HERE, "swe_napi.node")
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-DNODE_GYP_MODULE_NAME=swe_napi",
           "-DNAPI_VERSION=8", f"-I{NODE_INC}", "-o", out + ".tmp", os.path.join(HERE, "swe_napi.cc"),
           "-L" + os.path.dirname(HERE), "-lswe_amd", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("addon build failed")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build())
