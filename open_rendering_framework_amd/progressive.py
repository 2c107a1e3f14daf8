"""Progressive rendering of a scene file: the frame loop of OptixScene
(ORF/OptixScene.cpp:139-168 renderScene, :202-236 saveScene, :240-299
loadScene / loadBuffer) over the C-ABI.

    p = ProgressiveScene("scene.json", quit_and_save=True)
    while p.render_scene():   # one frame per call until frame_count frames
        pass                  # (saved as <scene>.json/.png/.raw at the end)

A scene whose document names a buffer_path that exists (and quit_and_save is
off) loads that film as the result of frame_count frames and renders nothing
more, as loadBuffer does.
"""
import os

from . import rt


class ProgressiveScene:
    def __init__(self, scene_path, quit_and_save=False, env_importance=False, ctx=None):
        self.ctx = ctx or rt.Context()
        self.scene_path = str(scene_path)
        self.quit_and_save = quit_and_save
        self.frame = 0
        self.max_frame = -1
        info = self.ctx.load_scene(self.scene_path, env_importance=env_importance)
        self.width, self.height = info["width"], info["height"]
        self.buffer_path = info["buffer_path"]
        self.frame = info["frame_count"]
        self.ctx.film_clear()
        self._load_buffer()

    def _load_buffer(self):
        """loadBuffer: an existing .raw is the film after frame_count frames."""
        if not self.quit_and_save and self.buffer_path and os.path.exists(self.buffer_path) and self.frame > 0:
            self.ctx.load_raw(self.buffer_path, self.width, self.height, self.frame)
            self.max_frame = self.frame
        elif self.frame > 0:
            self.max_frame = self.frame
            self.frame = 0

    def render_scene(self):
        """renderScene: renders the next frame if any is left; saves the scene
        after the last one when quit_and_save is on.  Returns True while frames remain."""
        if self.max_frame < 0 or self.frame < self.max_frame:
            self.ctx.render_frames(self.width, self.height, self.frame, self.frame + 1)
            self.frame += 1
            if self.frame == self.max_frame and self.quit_and_save:
                self.save_scene(self.scene_path)
        return self.max_frame < 0 or self.frame < self.max_frame

    def save_scene(self, path):
        """saveScene: <base>.json with frame_count and buffer_path, <base>.png, <base>.raw."""
        self.ctx.save_scene(path, self.frame)

    def image(self):
        """The film's mean, (height, width, 4) FP32, row 0 = the bottom of the image."""
        return self.ctx.film(self.width, self.height).astype("float32")
