"""ORACLE — test infrastructure only: CPU restatements of the reference
algorithms used to check libnlod.  The product never imports this package."""
