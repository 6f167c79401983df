from .sigmas import *  # noqa: F401,F403
