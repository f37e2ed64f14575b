"""Package logger (the reference uses logging.getLogger('decontamination'), decontamination/logging.py:15)."""
import logging as _logging

logger = _logging.getLogger('decontamination')
