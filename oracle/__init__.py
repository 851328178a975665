"""CPU oracle (test infrastructure only): a C restatement of the reference's LJ force and
integration path, loaded through oracle.pyoracle.  Never imported by the product."""
