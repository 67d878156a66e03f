"""Package and paper-artifact version (torch_sbl/version.py:1-4).

The release identifier names the numerical core the locked profiles are
defined against; the profiles check it before building a plan.
"""

__version__ = "1.1.0"
PAPER_RELEASE_ID = "FBME-SBL-TAES-v1.1"
