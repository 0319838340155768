"""Writes configuration_vectors.json: the golden values of the reference's
ConfigurationTest.java (src/test/java/br/unb/cic/tdp/base/ConfigurationTest.java,
configurations 1-3: signature :33-40, canonical signature :42-49, hash code
:51-55, equivalent signatures :57-92, and the same for configurations 2 and 3).
Run in a checkout that has the reference; the JSON is the committed fixture."""
import json
import os
import re
import sys

SRC = os.environ.get("TDP_REFERENCE", "/root/reference")
PATH = os.path.join(SRC, "src/test/java/br/unb/cic/tdp/base/ConfigurationTest.java")


def floats(s):
    return [float(x.strip().rstrip("f")) for x in s.split(",") if x.strip()]


def main(out):
    text = open(PATH).read()
    configs = {}
    for m in re.finditer(r"void testConfig(\d)_(\w+)\(\)\s*\{(.*?)\n    \}", text, re.S):
        k, kind, body = m.group(1), m.group(2), m.group(3)
        cfg = configs.setdefault(k, {})
        cfg["spi"] = re.search(r'new Configuration\("([^"]+)"\)', body).group(1)
        if kind == "Signature":
            cfg["signature"] = floats(re.search(r"expectedSignature = \{([^}]*)\}", body).group(1))
        elif kind == "CanonicalSignature":
            cfg["canonical"] = floats(re.search(r"expectedCanonicalSignature = \{([^}]*)\}", body).group(1))
        elif kind == "HashCode":
            cfg["hash"] = int(re.search(r"assertEquals\((-?\d+), config\.hashCode\(\)\)", body).group(1))
        elif kind == "EquivalentSignatures":
            cfg["equivalent"] = [floats(s) for s in re.findall(r'"\[([^\]]*)\]"', body)]
    data = {"source": "ConfigurationTest.java (reference), configurations 1-3",
            "configurations": [configs[k] for k in sorted(configs)]}
    with open(out, "w") as fh:
        json.dump(data, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "configuration_vectors.json"))
