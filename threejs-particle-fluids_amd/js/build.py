"""Build the Node N-API addon pf_node.node against libpf.so (g++, node headers)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
NODE_INC = "/usr/include/node"


def build():
    out = os.path.join(HERE, "pf_node.node")
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I" + NODE_INC, "-DNODE_GYP_MODULE_NAME=pf_node",
           os.path.join(HERE, "addon.cc"), "-o", out + ".tmp", "-L" + os.path.dirname(HERE), "-lpf",
           "-Wl,-rpath,$ORIGIN/.."]
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    if not os.path.isdir(NODE_INC):
        sys.exit("node headers not found")
    print(build())
