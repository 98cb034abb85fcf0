"""Writes tests/golden/temoa-g1-equil-excerpt.xml: the reference's equilibrated
State checkpoint (python/tests/temoa-g1-equil.xml) cut to its first 16
particles -- root attributes, box, the Parameters element with the ATM global
parameters, 16 Position and 16 Velocity elements, IntegratorParameters --
byte for byte as they appear in the original file.  Run in a tree where the
reference checkout is present:  python tests/golden/make_state_excerpt.py
"""
import os
import re
import sys

SRC = "/root/reference/python/tests/temoa-g1-equil.xml"
KEEP = 16


def main():
    text = open(SRC).read()
    head = text[:text.index("\t<Positions>")]
    pos = re.findall(r"\t\t<Position [^>]*/>\n", text)
    vel = re.findall(r"\t\t<Velocity [^>]*/>\n", text)
    tail = text[text.index("\t<IntegratorParameters"):]
    out = head + "\t<Positions>\n" + "".join(pos[:KEEP]) + "\t</Positions>\n" + \
        "\t<Velocities>\n" + "".join(vel[:KEEP]) + "\t</Velocities>\n" + tail
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "temoa-g1-equil-excerpt.xml")
    open(dst, "w").write(out)
    print(f"{dst}: {len(pos)} -> {KEEP} particles", file=sys.stderr)


if __name__ == "__main__":
    main()
