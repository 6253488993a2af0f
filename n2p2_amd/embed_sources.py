"""Writes the device headers the run-time (hipRTC) build compiles as C++
raw-string literals: python3 embed_sources.py OUT HEADER..."""
import os
import sys

DELIM = "HDNNP_RTC_SRC"


def main():
    out, headers = sys.argv[1], sys.argv[2:]
    names, bodies = [], []
    for h in headers:
        text = open(h).read()
        if (")" + DELIM + '"') in text:
            raise SystemExit("delimiter found in " + h)
        names.append(os.path.basename(h))
        bodies.append('R"%s(%s)%s"' % (DELIM, text, DELIM))
    with open(out, "w") as f:
        f.write("static const int kRtcNumHeaders = %d;\n" % len(headers))
        f.write("static const char* kRtcHeaderNames[] = {%s};\n"
                % ", ".join('"%s"' % n for n in names))
        f.write("static const char* kRtcHeaders[] = {\n%s};\n"
                % ",\n".join(bodies))


if __name__ == "__main__":
    main()
