"""Writes extension_roots.json: the roots of the two case searches and the
canonical configurations the reference's proof pages name for them.

* Extensions.java:41-46 (DFS roots) and Combinations.java:31-42 (bad small
  components): the configurations as the code builds them.
* index.html:44-49,54-56 link every DFS root and bad small component to
  ``dfs/<canonical spi>.html`` (SortOrExtend names a configuration's pages and
  bad-case files by ``getCanonical().getSpi()``, SortOrExtend.java:55-75), and
  index.html:66-71 link the combination roots by their own spi; the link text
  pairs the two.
* index.html:59: the full configuration the DFS reached that does not allow
  an 11/8-sequence.

Run in a checkout that has the reference; the JSON is the committed fixture."""
import json
import os
import re
import sys

SRC = os.environ.get("TDP_REFERENCE", "/root/reference")
JAVA = os.path.join(SRC, "src/main/java/br/unb/cic/tdp/proof/seq11_8")
HTML = os.path.join(SRC, "src/main/resources/index.html")


def java_roots(name, pattern):
    text = open(os.path.join(JAVA, name)).read()
    return re.findall(pattern, text)


def main(out):
    html = open(HTML).read()
    links = re.findall(r'<a href="(dfs|comb)/([^"]+)\.html">([^<]+)</a>', html)
    dfs = {}
    comb = {}
    for kind, spi, text in links:
        (dfs if kind == "dfs" else comb).setdefault(text.strip(), spi)
    dfs_roots = java_roots("Extensions.java", r'new MulticyclePermutation\("([^"]+)"\)\), outputDir \+ "/dfs/"\)\);\s*\n')
    components = java_roots("Combinations.java", r'new MulticyclePermutation\(\s*"([^"]+)"\)\);')
    full = re.search(r"The search found the full configuration <a[^>]*>([^<]+)</a>", html).group(1)
    data = {
        "dfs_roots": dfs_roots,
        "bad_small_components": components,
        # link text -> canonical spi (dfs pages) and own spi (comb pages)
        "dfs_canonical_by_name": dfs,
        "comb_spi_by_name": comb,
        "full_configuration_without_11_8": full,
    }
    with open(out, "w") as fh:
        json.dump(data, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "extension_roots.json"))
