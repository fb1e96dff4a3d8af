"""The ``pyop2`` logger (reference: pyop2/logger.py:36-53)."""

import logging

logger = logging.getLogger('pyop2')
_handler = logging.StreamHandler()
_handler.setFormatter(logging.Formatter('%(name)s:%(levelname)s %(message)s'))
logger.addHandler(_handler)
logger.setLevel(logging.WARNING)

debug = logger.debug
info = logger.info
warning = logger.warning
error = logger.error
critical = logger.critical


def set_log_level(level):
    """Set the log level of the pyop2 logger ('WARN' is accepted)."""
    if level == 'WARN':
        level = 'WARNING'
    logger.setLevel(level)
