"""
Utility Module (:mod:`vulkpy_amd.util`)
=======================================

Same entry points as vulkpy/util.py (enable_debug 19-55, getShader 58-72).

Examples
--------
>>> from vulkpy_amd.util import enable_debug
>>> enable_debug(api_dump=False)
"""

import logging
import os

logger = logging.getLogger("vulkpy")


def enable_debug(*, validation: bool = True, api_dump: bool = True):
    """
    Enable debug message

    Parameters
    ----------
    validation : bool, optional
        If ``True`` (default), make every HIP launch synchronous
        (``HIP_LAUNCH_BLOCKING=1``) so that an error is reported by the call
        that caused it.  Takes effect for contexts created afterwards.
    api_dump : bool, optional
        If ``True`` (default), log HIP API calls (``AMD_LOG_LEVEL=3``).

    Notes
    -----
    The reference enables the Vulkan validation and API-dump layers here
    (vulkpy/util.py:43-55); these are the HIP runtime's counterparts.
    """
    logging.basicConfig(level=logging.DEBUG)
    logger.setLevel(logging.DEBUG)
    logger.debug("Enable debug mode")

    if validation:
        os.environ["HIP_LAUNCH_BLOCKING"] = "1"
        logger.debug("Enable synchronous HIP launches")
    if api_dump:
        os.environ["AMD_LOG_LEVEL"] = "3"
        logger.debug("Enable HIP API logging")


def getShader(name: str) -> str:
    """
    Name of the kernel family that replaces a reference shader

    Parameters
    ----------
    name : str
        SPIR-V (.spv) name of the reference

    Returns
    -------
    str
        The same name (kernels are compiled into ``libvulkpy_hip.so``; there
        are no shader files to locate).
    """
    return name
